"""Host test of the XCD-aware block order of the traversal kernels.

xcd_block (smallupbp_amd/csrc/traverse.hip) renumbers workgroups so that each of the 8 XCDs,
which the hardware deals consecutive workgroups to round-robin, works on one contiguous range
of tiles. Every tile must still be processed exactly once: the mapping has to be a bijection of
[0, n) for every grid size n, including sizes that are not multiples of 8. The test compiles
the function's own source text on the host and checks that for every n up to 4,100 and for the
grid sizes of the benchmark workloads, and that each XCD's tiles are contiguous."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdint>
#include <cstdio>
#include <vector>
#define __device__
#define __forceinline__ inline
%s
static int check(int64_t n) {
    std::vector<int> seen(n, 0);
    std::vector<int64_t> lo(8, INT64_MAX), hi(8, -1), cnt(8, 0);
    for (int64_t b = 0; b < n; ++b) {
        const int64_t t = xcd_block(b, n);
        if (t < 0 || t >= n || seen[t]++) return 1;
        const int x = (int)(b %% 8);
        if (t < lo[x]) lo[x] = t;
        if (t > hi[x]) hi[x] = t;
        ++cnt[x];
    }
    for (int x = 0; x < 8; ++x)
        if (cnt[x] && hi[x] - lo[x] + 1 != cnt[x]) return 2;  // one contiguous range per XCD
    return 0;
}
int main() {
    int bad = 0;
    for (int64_t n = 1; n <= 4100; ++n)
        if (int e = check(n)) { std::printf("n=%%lld error %%d\n", (long long)n, e); ++bad; }
    const int64_t big[] = {32400, 16200, 8100, 15625, 1000192 / 256, 2073600 / 64, 2073600 / 128, 999999};
    for (int64_t n : big)
        if (int e = check(n)) { std::printf("n=%%lld error %%d\n", (long long)n, e); ++bad; }
    return bad;
}
"""


def test_xcd_block_is_a_bijection_with_contiguous_xcd_ranges(tmp_path):
    src = open(os.path.join(ROOT, "smallupbp_amd", "csrc", "traverse.hip")).read()
    m = re.search(r"__device__ __forceinline__ int64_t xcd_block\(int64_t bid, int64_t nblk\) \{.*?\n\}", src, re.S)
    assert m, "xcd_block not found in traverse.hip"
    cpp = tmp_path / "xcd.cpp"
    cpp.write_text(HARNESS % m.group(0))
    exe = tmp_path / "xcd"
    try:
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", str(cpp), "-o", str(exe)])
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.fail("host build of xcd_block failed: %s" % e)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
