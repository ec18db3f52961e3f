"""Row-legal synthetic designs for the detailed-placement tests (test helper, not a test module).

abutted_design lays standard cells left to right along each row, in runs of abutting cells
separated by gaps of a few sites (so each row holds several segments), and adds fixed 1 x 1
terminals on the core boundary.  With private=True every net connects cells of one group of four
consecutive cells of a run (groups start at the run's first cell, the windows of a phase-0 pass)
plus, with probability terminal_share, a terminal: within one pass no two windows then share a
net.  Otherwise nets draw their pins uniformly from all cells and terminals with the degree mix
of the synthetic netlists (2: 60 %, 3: 25 %, 4-8: 10 %, 9-30: 5 %).
"""

import numpy as np

from myplacement_amd import design


def abutted_design(seed: int = 1, rows: int = 12, sites: int = 160, row_height: float = 9.0, private: bool = False,
                   nets_per_cell: float = 1.1, terminals: int = 12, fill: float = 0.85,
                   terminal_share: float = 0.3) -> design.Design:
    rng = np.random.default_rng(seed)
    widths, centers, runs = [], [], []
    for r in range(rows):
        x = float(rng.integers(0, 4))
        run = []
        while True:
            w = float(rng.integers(1, 9))
            if x + w > sites:
                break
            if run and rng.random() > fill:  # end the run: leave a gap of 1-5 sites
                runs.append(run)
                run = []
                x += float(rng.integers(1, 6))
                if x + w > sites:
                    break
            run.append(len(widths))
            widths.append(w)
            centers.append((x + w / 2, r * row_height + row_height / 2))
            x += w
        if run:
            runs.append(run)
    cells = len(widths)
    W, H = float(sites), rows * row_height
    tpos = [(W * (k + 0.5) / terminals, -0.5 if k % 2 == 0 else H + 0.5) for k in range(terminals)]
    width = np.array(widths + [1.0] * terminals)
    height = np.array([row_height] * cells + [1.0] * terminals)
    center = np.array(centers + tpos)
    is_term = np.array([0] * cells + [1] * terminals, dtype=np.uint8)
    nets = []
    if private:
        for run in runs:
            for g in range(0, len(run), 4):
                group = run[g:g + 4]
                if len(group) < 2:
                    continue
                for _ in range(int(rng.integers(1, 4))):
                    k = int(rng.integers(2, len(group) + 1))
                    pins = list(rng.choice(group, size=k, replace=False))
                    if rng.random() < terminal_share:
                        pins.append(cells + int(rng.integers(0, terminals)))
                    nets.append(pins)
    else:
        hosts = cells + terminals
        for _ in range(int(nets_per_cell * cells)):
            u = rng.random()
            deg = 2 if u < 0.6 else 3 if u < 0.85 else int(rng.integers(4, 9)) if u < 0.95 else int(rng.integers(9, 31))
            nets.append(list(rng.choice(hosts, size=min(deg, hosts), replace=False)))
    start = np.zeros(len(nets) + 1, dtype=np.int64)
    pin_module, pin_offset = [], []
    for n, pins in enumerate(nets):
        for m in pins:
            pin_module.append(int(m))
            pin_offset.append(((rng.random() - 0.5) * width[m], (rng.random() - 0.5) * height[m]))
        start[n + 1] = len(pin_module)
    row_arr = np.array([(r * row_height, row_height, 0.0, 1.0, float(sites)) for r in range(rows)])
    M = cells + terminals
    return design.Design(width=width, height=height, center=center, orientation=np.zeros(M, dtype=np.int32),
                         is_terminal=is_term, is_fixed=is_term.copy(), net_weight=np.ones(len(nets)),
                         net_pin_start=start, pin_module=np.array(pin_module, dtype=np.int32),
                         pin_offset=np.array(pin_offset).reshape(-1, 2), rows=row_arr)


def reordering_fixture() -> design.Design:
    """tests/DetailedPlacementTests.cpp:14-43: cells A, C, D, B (5 x 10) abutting from x = 0 in one
    40-site row, nets A-B and C-D; one four-cell reorder takes HPWL from 20 to 10."""
    ll = np.array([0.0, 5.0, 10.0, 15.0])
    return design.Design(width=np.full(4, 5.0), height=np.full(4, 10.0),
                         center=np.stack([ll + 2.5, np.full(4, 5.0)], 1), orientation=np.zeros(4, dtype=np.int32),
                         is_terminal=np.zeros(4, dtype=np.uint8), is_fixed=np.zeros(4, dtype=np.uint8),
                         net_weight=np.ones(2), net_pin_start=np.array([0, 2, 4]),
                         pin_module=np.array([0, 3, 1, 2], dtype=np.int32), pin_offset=np.zeros((4, 2)),
                         rows=np.array([[0.0, 10.0, 0.0, 1.0, 40.0]]))
