"""Host regression test of the 64-bit wave-uniform widening in the traversal kernels.

uni_u64 (smallupbp_amd/csrc/traverse.hip) makes a 64-bit value wave-uniform through two 32-bit
readfirstlane calls, whose builtin returns int. Widening the low half through int instead of
uint32_t sign-extends a low word with bit 31 set over the high word (a wrong device address;
this once faulted the GPU). The test compiles the function's own source text on the host with
readfirstlane as the identity (what it is for a wave-uniform value) and checks the round trip
on values whose low and high words have bit 31 set and clear."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdint>
#include <cstdio>
#define __device__
#define __forceinline__ inline
static int readfirstlane_identity(int v) { return v; }
#define __builtin_amdgcn_readfirstlane readfirstlane_identity
%s
int main() {
    const uint64_t vals[] = {0ull, 1ull, 0x80000000ull, 0xFFFFFFFFull, 0x100000000ull, 0x7FFFFFFF80000000ull,
                             0x00007F3A9C000000ull | 0x80000000ull, 0xFFFFFFFF00000000ull, 0xFFFFFFFFFFFFFFFFull,
                             0x123456789ABCDEF0ull, 0x00007FFF_FFFFFFFFull};
    int bad = 0;
    for (uint64_t v : vals)
        if (uni_u64(v) != v) {
            std::printf("uni_u64(%%016llx) = %%016llx\n", (unsigned long long)v, (unsigned long long)uni_u64(v));
            ++bad;
        }
    return bad;
}
"""


def test_uni_u64_round_trips_values_with_bit31_set(tmp_path):
    src = open(os.path.join(ROOT, "smallupbp_amd", "csrc", "traverse.hip")).read()
    m = re.search(r"__device__ __forceinline__ uint64_t uni_u64\(uint64_t v\) \{.*?\n\}", src, re.S)
    assert m, "uni_u64 not found in traverse.hip"
    cpp = tmp_path / "widen.cpp"
    cpp.write_text((HARNESS % m.group(0)).replace("_FFFFFFFF", "FFFFFFFF"))
    exe = tmp_path / "widen"
    try:
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", str(cpp), "-o", str(exe)])
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.fail("host build of uni_u64 failed: %s" % e)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
