"""The onesweep look-back word (palzilla_amd/csrc/lookback.hpp) on the CPU:
the header is compiled with g++ into a small checker and run.  It packs a
tag (sort epoch * 8 + pass), a flag and a 30-bit count; a word is current only
for its own tag.  Checked at the ends of every field and at the epochs where
a 30-bit tag field would have wrapped (2^27 sorts) and where the host clears
the table (2^28)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "palzilla_amd", "csrc")

PROGRAM = r"""
#include <cstdio>
#include <initializer_list>
#include "lookback.hpp"
using namespace palz;
static int bad = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL line %d: %s\n", __LINE__, #c); ++bad; } } while (0)
int main() {
    const uint32_t epochs[] = {1u, 2u, (1u << 27) - 1u, 1u << 27, (1u << 27) + 1u, 1u << 28};
    const uint32_t counts[] = {0u, 1u, 4096u, kLbCountMask};
    for (uint32_t e : epochs)
        for (uint32_t pass = 0; pass < kLbPassesPerSort; ++pass)
            for (uint32_t c : counts)
                for (uint32_t f : {kLbFlagAggregate, kLbFlagInclusive}) {
                    const uint32_t tag = lb_tag(e, pass);
                    const uint64_t w = lb_pack(tag, f, c);
                    CHECK(lb_current(w, tag));
                    CHECK(lb_flag(w) == f);
                    CHECK(lb_count(w) == c);
                    // another pass, the previous sort and the sort 2^27 earlier never match
                    CHECK(!lb_current(w, lb_tag(e, (pass + 1) % kLbPassesPerSort)));
                    CHECK(!lb_current(w, lb_tag(e - 1u, pass)));
                    if (e > (1u << 27)) CHECK(!lb_current(w, lb_tag(e - (1u << 27), pass)));
                }
    // a cleared word (zero) matches no tag a sort uses: epochs start at 1
    CHECK(!lb_current(0ull, lb_tag(1u, 0u)));
    // the tags of epochs 1 .. 2^28 are distinct and below 2^32
    CHECK(lb_tag((uint32_t)kLbSortsPerClear, kLbPassesPerSort - 1) > lb_tag((uint32_t)kLbSortsPerClear - 1u, 0u));
    CHECK((uint64_t)kLbSortsPerClear * kLbPassesPerSort + kLbPassesPerSort < (1ull << 32));
    std::printf(bad ? "%d failures\n" : "ok\n", bad);
    return bad != 0;
}
"""


def test_lookback_word_fields_and_tags(tmp_path):
    src = tmp_path / "lb.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "lb"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", HDR, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip() == "ok"
