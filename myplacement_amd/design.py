"""Netlist container, BookShelf reader and the seeded synthetic designs.

`Design` holds the PlacementDatabase fields the global-placement path reads
(include/myplacement/model/PlacementDatabase.hpp:25-71) as flat numpy arrays, with pins in net
order (the pin-id order src/io/BookshelfParser.cpp:244-271 produces).

`read_bookshelf` follows the record rules of src/io/BookshelfParser.cpp:62-348 (.aux/.nodes/.scl/
.pl/.nets/.wts, declared-count checks).  `synthetic_design` builds the benchmark designs from a
seed; their exact recipe is written in each builder's docstring.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

ORIENTATIONS = ("N", "S", "E", "W", "FN", "FS", "FE", "FW")


class ParseError(RuntimeError):
    """Mirrors myplacement::ParseError (include/myplacement/io/BookshelfParser.hpp)."""


@dataclass
class Design:
    width: np.ndarray            # [M] f64
    height: np.ndarray           # [M] f64
    center: np.ndarray           # [M,2] f64
    orientation: np.ndarray      # [M] i32
    is_terminal: np.ndarray      # [M] u8
    is_fixed: np.ndarray         # [M] u8
    net_weight: np.ndarray       # [N] f64
    net_pin_start: np.ndarray    # [N+1] i64
    pin_module: np.ndarray       # [I] i32
    pin_offset: np.ndarray       # [I,2] f64
    rows: np.ndarray             # [R,5] f64: bottom, height, x_start, site_spacing, site_count
    names: list | None = None
    net_names: list | None = None
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.width = np.ascontiguousarray(self.width, dtype=np.float64)
        self.height = np.ascontiguousarray(self.height, dtype=np.float64)
        self.center = np.ascontiguousarray(self.center, dtype=np.float64).reshape(-1, 2)
        self.orientation = np.ascontiguousarray(self.orientation, dtype=np.int32)
        self.is_terminal = np.ascontiguousarray(self.is_terminal, dtype=np.uint8)
        self.is_fixed = np.ascontiguousarray(self.is_fixed, dtype=np.uint8)
        self.net_weight = np.ascontiguousarray(self.net_weight, dtype=np.float64)
        self.net_pin_start = np.ascontiguousarray(self.net_pin_start, dtype=np.int64)
        self.pin_module = np.ascontiguousarray(self.pin_module, dtype=np.int32)
        self.pin_offset = np.ascontiguousarray(self.pin_offset, dtype=np.float64).reshape(-1, 2)
        self.rows = np.ascontiguousarray(self.rows, dtype=np.float64).reshape(-1, 5)

    @property
    def module_count(self) -> int:
        return int(self.width.shape[0])

    @property
    def net_count(self) -> int:
        return int(self.net_weight.shape[0])

    @property
    def pin_count(self) -> int:
        return int(self.pin_module.shape[0])

    def core_region(self) -> tuple:
        """Bounding box of the rows (PlacementDatabase.cpp:170-175)."""
        r = self.rows
        x_end = r[:, 2] + r[:, 4] * r[:, 3]
        return (float(r[:, 2].min()), float(r[:, 0].min()), float(x_end.max()), float((r[:, 0] + r[:, 1]).max()))

    def to_c(self, center: np.ndarray | None = None, orientation: np.ndarray | None = None) -> abi.MpDesign:
        """A C view of this design (with other centres / orientations if given); the arrays stay
        owned by `self`."""
        c = self.center if center is None else np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 2)
        o = self.orientation if orientation is None else np.ascontiguousarray(orientation, dtype=np.int32)
        self._keep = [c, o]
        return abi.MpDesign(
            module_count=self.module_count,
            width=abi.dptr(self.width), height=abi.dptr(self.height), center=abi.dptr(c),
            orientation=abi.i32ptr(o), is_terminal=abi.u8ptr(self.is_terminal),
            is_fixed=abi.u8ptr(self.is_fixed), net_count=self.net_count,
            net_weight=abi.dptr(self.net_weight), net_pin_start=abi.i64ptr(self.net_pin_start),
            pin_count=self.pin_count, pin_module=abi.i32ptr(self.pin_module),
            pin_offset=abi.dptr(self.pin_offset), row_count=int(self.rows.shape[0]), rows=abi.dptr(self.rows),
        )

    def movable_ids(self) -> np.ndarray:
        fixed = (self.is_fixed != 0) | (self.is_terminal != 0)
        return np.nonzero(~fixed)[0].astype(np.int64)


# ---------------------------------------------------------------------------------------------
# BookShelf reader (src/io/BookshelfParser.cpp)

def _strip(line: str) -> str:
    return line.split("#", 1)[0].strip()


def _numbers(line: str) -> list:
    out = []
    for tok in line.replace(":", " ").replace(",", " ").replace("\t", " ").split():
        try:
            out.append(float(tok))
        except ValueError:
            pass
    return out


def read_bookshelf(aux_path: str) -> Design:
    """BookshelfParser::parseAux (BookshelfParser.cpp:313-348)."""
    directory = os.path.dirname(aux_path)
    files = {}
    with open(aux_path) as fh:
        for raw in fh:
            line = _strip(raw)
            if not line or ":" not in line:
                continue
            for name in line.split(":", 1)[1].split():
                ext = os.path.splitext(name)[1]
                if ext and ext not in files:
                    files[ext] = os.path.join(directory, name)
    for ext in (".nodes", ".scl", ".pl", ".nets"):
        if ext not in files:
            raise ParseError(f"{aux_path}: missing required '{ext}' file in AUX record")

    names, width, height, terminal = [], [], [], []
    index = {}
    declared = {}
    with open(files[".nodes"]) as fh:  # :62-99
        for n, raw in enumerate(fh, 1):
            line = _strip(raw)
            if not line or line.startswith("UCLA"):
                continue
            nums = _numbers(line)
            if line.startswith("NumNodes") or line.startswith("NumTerminals"):
                if not nums:
                    raise ParseError(f"{files['.nodes']}: invalid {line.split()[0]} at line {n}")
                declared[line.split()[0].rstrip(":")] = int(nums[0])
                continue
            tok = line.split()
            if len(tok) < 3:
                raise ParseError(f"{files['.nodes']}: invalid node at line {n}")
            try:
                w, h = float(tok[1]), float(tok[2])
            except ValueError as exc:
                raise ParseError(f"{files['.nodes']}: invalid node at line {n}") from exc
            if tok[0] in index:
                raise ValueError(f"Duplicate module name: {tok[0]}")
            if w < 0 or h < 0:
                raise ValueError(f"A module cannot have a negative size: {tok[0]}")
            index[tok[0]] = len(names)
            names.append(tok[0])
            width.append(w)
            height.append(h)
            terminal.append(any("TERMINAL" in a.upper() for a in tok[3:]))
    M = len(names)
    fixed = list(terminal)
    center = np.zeros((M, 2))
    orient = np.zeros(M, dtype=np.int32)

    rows = []
    inside = False
    with open(files[".scl"]) as fh:  # :101-164
        for n, raw in enumerate(fh, 1):
            line = _strip(raw)
            if not line or line.startswith("UCLA") or line.startswith("NumRows"):
                continue
            if line.startswith("CoreRow"):
                inside = True
                bottom, rh, spacing = 0.0, 0.0, 1.0
                continue
            if not inside:
                continue
            if line == "End":
                inside = False
                continue
            nums = _numbers(line)
            if line.startswith("Coordinate") and nums:
                bottom = nums[0]
            elif line.startswith("Height") and nums:
                rh = nums[0]
            elif line.startswith("Sitespacing") and nums:
                spacing = nums[0]
            elif line.startswith("SubrowOrigin"):
                if len(nums) < 2 or rh <= 1e-9 or spacing <= 1e-9:
                    raise ParseError(f"{files['.scl']}: invalid SubrowOrigin at line {n}")
                rows.append((bottom, rh, nums[0], spacing, float(int(nums[1]))))
    if not rows:
        raise ParseError(f"{files['.scl']}: contains no valid SiteRow records")

    with open(files[".pl"]) as fh:  # :166-202
        for n, raw in enumerate(fh, 1):
            line = _strip(raw)
            if not line or line.startswith("UCLA"):
                continue
            tok = line.split()
            if len(tok) < 3:
                continue
            try:
                x, y = float(tok[1]), float(tok[2])
            except ValueError:
                continue
            if tok[0] not in index:
                raise ParseError(f"{files['.pl']}: position references unknown module '{tok[0]}' at line {n}")
            m = index[tok[0]]
            rest = tok[3:]
            if rest:
                o = rest[0]
                if o == ":":
                    o = rest[1] if len(rest) > 1 else ""
                    rest = rest[2:]
                else:
                    o = o[1:] if o.startswith(":") else o
                    rest = rest[1:]
                if o:
                    orient[m] = ORIENTATIONS.index(o.upper()) if o.upper() in ORIENTATIONS else 8
            if any("FIXED" in a.upper() for a in rest):
                fixed[m] = True
            center[m] = (x + width[m] * 0.5, y + height[m] * 0.5)

    net_names, net_start, pin_module, pin_offset = [], [0], [], []
    declared_nets = declared_pins = None
    with open(files[".nets"]) as fh:  # :204-273
        lines = [(_strip(raw), n) for n, raw in enumerate(fh, 1)]
    i = 0
    while i < len(lines):
        line, n = lines[i]
        i += 1
        if not line or line.startswith("UCLA"):
            continue
        nums = _numbers(line)
        if line.startswith("NumNets"):
            declared_nets = int(nums[0]) if nums else None
            continue
        if line.startswith("NumPins"):
            declared_pins = int(nums[0]) if nums else None
            continue
        if not line.startswith("NetDegree"):
            continue
        tok = line.split()
        if len(tok) < 4:
            raise ParseError(f"{files['.nets']}: invalid NetDegree at line {n}")
        degree, name = int(tok[2]), tok[3]
        net_names.append(name)
        for _ in range(degree):
            while i < len(lines) and not lines[i][0]:
                i += 1
            if i >= len(lines):
                raise ParseError(f"{files['.nets']}: unexpected end of file while reading net '{name}'")
            pl, pn = lines[i]
            i += 1
            ptok = pl.replace(":", " ").split()
            if len(ptok) < 4:
                raise ParseError(f"{files['.nets']}: invalid pin at line {pn}")
            if ptok[0] not in index:
                raise ParseError(f"{files['.nets']}: net '{name}' references unknown module '{ptok[0]}' at line {pn}")
            pin_module.append(index[ptok[0]])
            pin_offset.append((float(ptok[2]), float(ptok[3])))
        net_start.append(len(pin_module))

    weights = np.ones(len(net_names))
    if ".wts" in files and os.path.exists(files[".wts"]):  # :275-288
        by_name = {nm: k for k, nm in enumerate(net_names)}
        with open(files[".wts"]) as fh:
            for raw in fh:
                line = _strip(raw)
                if not line or line.startswith("UCLA"):
                    continue
                tok = line.split()
                if len(tok) >= 2 and tok[0] in by_name:
                    try:
                        weights[by_name[tok[0]]] = float(tok[1])
                    except ValueError:
                        pass

    # validateCounts, :290-309
    if declared.get("NumNodes") is not None and declared["NumNodes"] != M:
        raise ParseError(f"{aux_path}: declared NumNodes={declared['NumNodes']}, parsed {M}")
    if declared.get("NumTerminals") is not None and declared["NumTerminals"] != sum(terminal):
        raise ParseError(f"{aux_path}: declared NumTerminals={declared['NumTerminals']}, parsed {sum(terminal)}")
    if declared_nets is not None and declared_nets != len(net_names):
        raise ParseError(f"{aux_path}: declared NumNets={declared_nets}, parsed {len(net_names)}")
    if declared_pins is not None and declared_pins != len(pin_module):
        raise ParseError(f"{aux_path}: declared NumPins={declared_pins}, parsed {len(pin_module)}")

    return Design(
        width=np.array(width), height=np.array(height), center=center, orientation=orient,
        is_terminal=np.array(terminal, dtype=np.uint8), is_fixed=np.array(fixed, dtype=np.uint8),
        net_weight=weights, net_pin_start=np.array(net_start, dtype=np.int64),
        pin_module=np.array(pin_module, dtype=np.int32),
        pin_offset=np.array(pin_offset, dtype=np.float64).reshape(-1, 2), rows=np.array(rows).reshape(-1, 5),
        names=names, net_names=net_names,
    )


# ---------------------------------------------------------------------------------------------
# Seeded synthetic designs

def _degrees(rng: np.random.Generator, n: int) -> np.ndarray:
    """Net degrees: 2 (60%), 3 (25%), U{4..8} (10%), U{9..30} (5%)."""
    u = rng.random(n)
    deg = np.where(u < 0.60, 2, np.where(u < 0.85, 3, 0))
    mid = (u >= 0.85) & (u < 0.95)
    deg[mid] = rng.integers(4, 9, size=int(mid.sum()))
    hi = u >= 0.95
    deg[hi] = rng.integers(9, 31, size=int(hi.sum()))
    return deg.astype(np.int64)


def _draw_pins(rng: np.random.Generator, degrees: np.ndarray, hosts: int) -> np.ndarray:
    """Distinct host nodes per net, uniform over `hosts` nodes (nets drawing a repeat are redrawn
    without replacement, in net order)."""
    total = int(degrees.sum())
    picks = rng.integers(0, hosts, size=total)
    net_of = np.repeat(np.arange(degrees.shape[0]), degrees)
    order = np.lexsort((picks, net_of))
    sp, sn = picks[order], net_of[order]
    dup = (sp[1:] == sp[:-1]) & (sn[1:] == sn[:-1])
    start = np.concatenate([[0], np.cumsum(degrees)])
    for k in np.unique(sn[1:][dup]):
        a, b = start[k], start[k + 1]
        picks[a:b] = rng.choice(hosts, size=b - a, replace=False)
    return picks


def _assemble(rng, width, height, center, is_term, is_fixed, degrees, hosts_idx, rows) -> Design:
    pin_host = _draw_pins(rng, degrees, hosts_idx.shape[0])
    pin_module = hosts_idx[pin_host].astype(np.int32)
    w = width[pin_module]
    h = height[pin_module]
    off = np.stack([(rng.random(pin_module.shape[0]) - 0.5) * w, (rng.random(pin_module.shape[0]) - 0.5) * h], axis=1)
    start = np.concatenate([[0], np.cumsum(degrees)]).astype(np.int64)
    M = width.shape[0]
    return Design(
        width=width, height=height, center=center, orientation=np.zeros(M, dtype=np.int32),
        is_terminal=is_term.astype(np.uint8), is_fixed=is_fixed.astype(np.uint8),
        net_weight=np.ones(degrees.shape[0]), net_pin_start=start, pin_module=pin_module,
        pin_offset=off, rows=rows,
    )


def plain_design(seed: int = 2026, cells: int = 2000, nets: int = 2300, row_count: int = 40,
                 sites: int = 400, row_height: float = 9.0, terminals: int = 64) -> Design:
    """The plain CPU-runnable case.

    `cells` standard cells of height `row_height` (one row) and integer width U{1..8} sites (site
    width 1); `row_count` rows of `sites` sites starting at (0, 0); `terminals` fixed 1x1 I/O
    terminals centred on the core boundary at equal perimeter spacing, starting at the lower-left
    corner and running counter-clockwise; `nets` nets with the degree mix of _degrees, pins drawn
    uniformly (distinct within a net) from cells plus terminals, pin offsets uniform within the
    node.  Movable cells start at the core centre (the random initial placement overwrites them).
    """
    rng = np.random.default_rng(seed)
    W, H = float(sites), row_count * row_height
    rows = np.array([(r * row_height, row_height, 0.0, 1.0, float(sites)) for r in range(row_count)])
    cw = rng.integers(1, 9, size=cells).astype(np.float64)
    ch = np.full(cells, row_height)
    perim = 2 * (W + H)
    tpos = []
    for k in range(terminals):
        s = (k + 0.5) * perim / terminals
        if s < W:
            tpos.append((s, 0.0))
        elif s < W + H:
            tpos.append((W, s - W))
        elif s < 2 * W + H:
            tpos.append((W - (s - W - H), H))
        else:
            tpos.append((0.0, H - (s - 2 * W - H)))
    width = np.concatenate([cw, np.ones(terminals)])
    height = np.concatenate([ch, np.ones(terminals)])
    center = np.concatenate([np.tile([W / 2, H / 2], (cells, 1)), np.array(tpos).reshape(-1, 2)])
    is_term = np.concatenate([np.zeros(cells), np.ones(terminals)])
    degrees = _degrees(rng, nets)
    hosts = np.arange(cells + terminals)
    return _assemble(rng, width, height, center, is_term, is_term, degrees, hosts, rows)


def large_design(seed: int = 2026, cells: int = 210_000, nets: int = 220_000, row_count: int = 1200,
                 sites: int = 1200, row_height: float = 9.0, macros: int = 12, macro_fraction: float = 0.15,
                 high_fanout_fraction: float = 0.002) -> Design:
    """The single large design at real shape.

    `cells` standard cells of height `row_height`, width = clip(round(exp(N(ln 3, 0.6))), 1, 40)
    sites (log-normal, median 3, max 40); `row_count` x `sites` core.  `macros` fixed rectangular
    macros (BookShelf terminals) covering `macro_fraction` of the core: equal area, width 1/4 of the
    core width, placed without overlap in two columns at seeded vertical positions.  `nets` nets
    with the degree mix of _degrees plus `high_fanout_fraction` of them at fanout U{50..500}; pins
    drawn uniformly from cells plus macros.  Movable cells start at the core centre.
    """
    rng = np.random.default_rng(seed)
    W, H = float(sites), row_count * row_height
    rows = np.array([(r * row_height, row_height, 0.0, 1.0, float(sites)) for r in range(row_count)])
    cw = np.clip(np.rint(np.exp(rng.normal(math.log(3.0), 0.6, size=cells))), 1, 40).astype(np.float64)
    ch = np.full(cells, row_height)
    mw = W / 4.0
    mh = macro_fraction * W * H / macros / mw
    per_col = (macros + 1) // 2
    slot = H / per_col
    mpos = []
    for k in range(macros):
        col, r = k % 2, k // 2
        x = (0.15 * W + mw / 2) if col == 0 else (0.60 * W + mw / 2)
        y = r * slot + mh / 2 + rng.random() * max(slot - mh, 0.0)
        mpos.append((x, y))
    width = np.concatenate([cw, np.full(macros, mw)])
    height = np.concatenate([ch, np.full(macros, mh)])
    center = np.concatenate([np.tile([W / 2, H / 2], (cells, 1)), np.array(mpos).reshape(-1, 2)])
    is_term = np.concatenate([np.zeros(cells), np.ones(macros)])
    degrees = _degrees(rng, nets)
    hf = rng.random(nets) < high_fanout_fraction
    degrees[hf] = rng.integers(50, 501, size=int(hf.sum()))
    hosts = np.arange(cells + macros)
    return _assemble(rng, width, height, center, is_term, is_term, degrees, hosts, rows)


def served_bins(cells: int) -> int:
    """Bin grid side of a served or batched job: the largest of 32, 64, 128, 256 that keeps at least
    4 cells per bin on average (32 when none does)."""
    side = 32
    for s_ in (64, 128, 256):
        if cells / (s_ * s_) >= 4.0:
            side = s_
    return side


def served_design(seed: int, cells: int) -> Design:
    """One job of the served mixed-size stream: the large design's recipe scaled to `cells` movable
    cells -- nets = 1.05 x cells, rows and sites scaled by sqrt(cells / 210,000) from 1,200 x 1,200
    (at least 40), max(1, round(12 x cells / 210,000)) macros covering 15 % of the core -- with a
    random initial placement from the same seed."""
    scale = math.sqrt(cells / 210_000.0)
    side = max(40, int(round(1200 * scale)))
    macros = max(1, int(round(12 * cells / 210_000)))
    d = large_design(seed=seed, cells=cells, nets=int(round(1.05 * cells)), row_count=side, sites=side,
                     macros=macros)
    return d


def batch_job_specs(count: int, seed: int = 2026) -> list:
    """The batching configuration's blocks: for block k, (job seed, cells, bin side) with cells
    log-uniform in 500-50k drawn from default_rng(seed) in block order, job seed = seed + 1 + k, and
    the bin side of served_bins.  Each block is served_design(job seed, cells) from
    a random initial placement with that seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        cells = int(round(math.exp(rng.uniform(math.log(500), math.log(50_000)))))
        out.append((seed + 1 + k, cells, served_bins(cells)))
    return out


def served_stream_specs(count: int, rate: float, seed: int = 2026) -> tuple:
    """The served stream's requests: for request k, (job seed, cells, bin side, target density) with
    cells log-uniform in 1k-200k and the target density uniform in 0.7-0.9, drawn in that order
    from default_rng(seed), job seed = seed + 1 + k; then the `count` Poisson inter-arrival gaps
    (exponential, mean 1 / rate) from the same generator.  Returns (specs, gaps)."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        cells = int(round(math.exp(rng.uniform(math.log(1_000), math.log(200_000)))))
        t = float(rng.uniform(0.7, 0.9))
        specs.append((seed + 1 + k, cells, served_bins(cells), t))
    gaps = rng.exponential(1.0 / rate, size=count)
    return specs, gaps


def _lower_left(c: float, half: float) -> float:
    """A lower-left coordinate ll with ll + half == c in FP64 (the reader forms the centre as
    lower-left + size / 2, PlacementDatabase.cpp Module::setLowerLeft), when one exists: c - half,
    else its nearest neighbours."""
    ll = c - half
    if ll + half == c:
        return ll
    up, down = ll, ll
    for _ in range(3):
        up = float(np.nextafter(up, np.inf))
        down = float(np.nextafter(down, -np.inf))
        if up + half == c:
            return up
        if down + half == c:
            return down
    return ll


def write_bookshelf(d: Design, directory: str, name: str) -> str:
    """Writes `d` as a BookShelf design (.aux/.nodes/.nets/.wts/.pl/.scl) that
    BookshelfParser::parseAux (BookshelfParser.cpp:62-348) and read_bookshelf read back into the same
    arrays.  Numbers are written with Python's shortest round-trip repr, so every size, offset,
    weight and row field comes back bit for bit; centres are written as lower-left corners chosen so
    that lower-left + size / 2 gives the centre back exactly when such a corner exists.  Modules and
    nets without names are called o<k> / n<k>.  Terminals are marked `terminal` in .nodes; other
    fixed modules `/FIXED` in .pl.  Returns the .aux path."""
    os.makedirs(directory, exist_ok=True)
    M, N = d.module_count, d.net_count
    names = d.names if d.names is not None else [f"o{k}" for k in range(M)]
    nets = d.net_names if d.net_names is not None else [f"n{k}" for k in range(N)]
    r = repr
    base = os.path.join(directory, name)
    with open(base + ".aux", "w") as f:
        f.write(f"RowBasedPlacement : {name}.nodes {name}.nets {name}.wts {name}.pl {name}.scl\n")
    with open(base + ".nodes", "w") as f:
        f.write("UCLA nodes 1.0\n\n")
        f.write(f"NumNodes : {M}\nNumTerminals : {int(d.is_terminal.sum())}\n")
        for k in range(M):
            f.write(f"{names[k]} {r(float(d.width[k]))} {r(float(d.height[k]))}"
                    + (" terminal\n" if d.is_terminal[k] else "\n"))
    with open(base + ".pl", "w") as f:
        f.write("UCLA pl 1.0\n\n")
        for k in range(M):
            o = int(d.orientation[k])
            ot = ORIENTATIONS[o] if 0 <= o < len(ORIENTATIONS) else "N"
            x = _lower_left(float(d.center[k, 0]), float(d.width[k]) * 0.5)
            y = _lower_left(float(d.center[k, 1]), float(d.height[k]) * 0.5)
            fixed = " /FIXED" if (d.is_fixed[k] and not d.is_terminal[k]) else ""
            f.write(f"{names[k]} {r(x)} {r(y)} : {ot}{fixed}\n")
    with open(base + ".nets", "w") as f:
        f.write("UCLA nets 1.0\n\n")
        f.write(f"NumNets : {N}\nNumPins : {d.pin_count}\n")
        for n in range(N):
            a, b = int(d.net_pin_start[n]), int(d.net_pin_start[n + 1])
            f.write(f"NetDegree : {b - a} {nets[n]}\n")
            for p in range(a, b):
                f.write(f"  {names[int(d.pin_module[p])]} I : {r(float(d.pin_offset[p, 0]))} "
                        f"{r(float(d.pin_offset[p, 1]))}\n")
    with open(base + ".wts", "w") as f:
        f.write("UCLA wts 1.0\n\n")
        for n in range(N):
            f.write(f"{nets[n]} {r(float(d.net_weight[n]))}\n")
    with open(base + ".scl", "w") as f:
        f.write("UCLA scl 1.0\n\n")
        f.write(f"NumRows : {d.rows.shape[0]}\n\n")
        for bottom, height, x0, spacing, count in d.rows:
            f.write("CoreRow Horizontal\n")
            f.write(f"  Coordinate : {r(float(bottom))}\n  Height : {r(float(height))}\n")
            f.write(f"  Sitewidth : {r(float(spacing))}\n  Sitespacing : {r(float(spacing))}\n")
            f.write("  Siteorient : N\n  Sitesymmetry : Y\n")
            f.write(f"  SubrowOrigin : {r(float(x0))} NumSites : {int(count)}\n")
            f.write("End\n")
    return base + ".aux"
