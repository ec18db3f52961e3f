"""The host set-up's stable radix sort (gp_host.cpp stableSortByKey) against std::stable_sort, no GPU
needed: keys with many ties over one, two and three 11-bit passes and inputs from empty to a million
items, under every host-thread limit.  The product's net order, particle renumbering and their
results depend on this order being exactly std::stable_sort's."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "myplacement_amd", "csrc")

PROGRAM = r"""
#include "gp_host.hpp"
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
int main() {
    std::mt19937 gen(7);
    int bad = 0;
    const int sizes[] = {0, 1, 1000, 70000, 300000, 1000003};
    const int bits[] = {5, 11, 20, 31};
    for (int limit : {1, 2, 3, 8}) {
        mpgp::t_parallel_limit = limit;
        for (int n : sizes)
            for (int b : bits) {
                std::vector<std::uint32_t> key(n), idx(n), want(n);
                const std::uint32_t mask = b >= 32 ? 0xffffffffU : ((1U << b) - 1U);
                for (int i = 0; i < n; ++i) key[i] = gen() & mask & 0x3ffU;  // ties everywhere
                if (b > 10)
                    for (int i = 0; i < n; i += 3) key[i] = gen() & mask;
                for (int i = 0; i < n; ++i) idx[i] = want[i] = static_cast<std::uint32_t>((i * 7919LL) % (n ? n : 1));
                std::stable_sort(want.begin(), want.end(), [&](std::uint32_t x, std::uint32_t y) { return key[x] < key[y]; });
                mpgp::stableSortByKey(idx.data(), idx.size(), key.data(), b);
                if (idx != want) {
                    std::printf("mismatch limit %d n %d bits %d\n", limit, n, b);
                    ++bad;
                }
            }
    }
    std::printf("%s\n", bad ? "FAIL" : "OK");
    return bad ? 1 : 0;
}
"""


def test_stable_radix_sort_equals_std_stable_sort(tmp_path):
    src = tmp_path / "sort_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "sort_check"
    r = subprocess.run(["g++", "-O2", "-std=c++20", "-pthread", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
                        str(src), os.path.join(CSRC, "gp_host.cpp"), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stdout + r.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout + out.stderr
