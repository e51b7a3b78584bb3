"""The launch-shape table (csrc/tune.cpp): parsing, validation, bucket choice and the text round trip, on the host.
tune.cpp is plain C++ (no HIP), so it is compiled here with a small driver; the GPU side (a context launching under a
table, rows identical under every table) is tests/test_gpu_tuning.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "postgis-cuda_amd", "csrc")

DRIVER = r"""
#include "tune.h"
#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
int main(int argc, char** argv) {
  std::stringstream in;
  in << std::cin.rdbuf();
  try {
    gsj::Tuning t = gsj::Tuning::parse(in.str(), "stdin");
    std::cout << t.text();
    for (int i = 1; i < argc; i++) {
      const gsj::LaunchShape& s = t.pick(std::stoull(argv[i]));
      std::cout << "pick " << argv[i] << " np>=" << s.np_min << " tcap=" << s.tcap << " edge_bpc=" << s.edge_bpc << "\n";
    }
    std::cout << "default " << gsj::Tuning{}.text();
    return 0;
  } catch (const std::invalid_argument& e) {
    std::cout << "ERR " << e.what() << "\n";
    return 3;
  }
}
"""


@pytest.fixture(scope="module")
def tunebin(tmp_path_factory):
    d = tmp_path_factory.mktemp("tune")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "tune_driver"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{CSRC}", str(src), os.path.join(CSRC, "tune.cpp"), "-ldl", "-o", str(exe)], check=True)
    return str(exe)


def run(exe, text, *picks):
    r = subprocess.run([exe, *map(str, picks)], input=text, capture_output=True, text=True)
    return r.returncode, r.stdout


def test_buckets_parse_pick_and_round_trip(tunebin):
    text = """# a comment line
bucket np>=0 tcap=0 edge_bpc=4          # small launches
bucket np>=32768 mcap=0 screen_occ=4
   bucket np>=2097152 edge_bpc=32 trav_block=128
"""
    rc, out = run(tunebin, text, 0, 32767, 32768, 2097151, 2097152, 10**12)
    assert rc == 0, out
    lines = out.splitlines()
    assert lines[0] == "# source: stdin"
    assert lines[1] == ("bucket np>=0 tcap=0 mcap=64 screen_occ=5 screen_bpc=16 edge_bpc=4 mixed_bpc=32 large_bpc=16 dw_bpc=16 "
                        "trav_block=256 ovf_lds=1 fork=1 mixed=1 large_cap=96 size_sort=1 dw_size_sort=0")
    assert lines[2].startswith("bucket np>=32768 tcap=16 mcap=0 screen_occ=4 ")
    assert lines[3].endswith("edge_bpc=32 mixed_bpc=32 large_bpc=16 dw_bpc=16 trav_block=128 ovf_lds=1 fork=1 mixed=1 large_cap=96 size_sort=1 dw_size_sort=0")
    picks = [l for l in lines if l.startswith("pick")]
    assert picks == ["pick 0 np>=0 tcap=0 edge_bpc=4", "pick 32767 np>=0 tcap=0 edge_bpc=4", "pick 32768 np>=32768 tcap=16 edge_bpc=16",
                     "pick 2097151 np>=32768 tcap=16 edge_bpc=16", "pick 2097152 np>=2097152 tcap=16 edge_bpc=32",
                     "pick 1000000000000 np>=2097152 tcap=16 edge_bpc=32"]
    # the printed table parses back to itself
    body = "\n".join(lines[1:4]) + "\n"
    rc2, out2 = run(tunebin, body)
    assert rc2 == 0 and out2.splitlines()[1:4] == lines[1:4]
    # the built-in table: one bucket of the shapes the kernels were written for
    assert [l for l in lines if l.startswith("default")] == ["default # source: built-in"]
    assert lines[-1] == ("bucket np>=0 tcap=16 mcap=64 screen_occ=5 screen_bpc=16 edge_bpc=16 mixed_bpc=32 large_bpc=16 dw_bpc=16 trav_block=256 "
                         "ovf_lds=1 fork=1 mixed=1 large_cap=96 size_sort=1 dw_size_sort=0")


@pytest.mark.parametrize("text,msg", [
    ("", "no bucket"),
    ("# only comments\n", "no bucket"),
    ("bucket tcap=16\n", "missing np>=N"),
    ("bucket np>=5\n", "must start at np>=0"),
    ("bucket np>=0\nbucket np>=0\n", "ascend"),
    ("bucket np>=0\nbucket np>=100\nbucket np>=50\n", "ascend"),
    ("bucket np>=0 tcap=17\n", "tcap outside [0, 16]"),
    ("bucket np>=0 mcap=32\n", "mcap must be 0 or above"),
    ("bucket np>=0 mcap=65\n", "mcap outside"),
    ("bucket np>=0 screen_occ=3\n", "screen_occ outside [4, 5]"),
    ("bucket np>=0 edge_bpc=0\n", "edge_bpc outside [1, 64]"),
    ("bucket np>=0 ovf_lds=3\n", "ovf_lds outside [0, 2]"),
    ("bucket np>=0 trav_block=96\n", "trav_block must be 64, 128 or 256"),
    ("bucket np>=0 warp=32\n", "unknown key 'warp'"),
    ("bucket np>=0 tcap=-1\n", "bad value"),
    ("bucket np>=0 tcap\n", "expected key=value"),
    ("bucket np>=x\n", "bad bucket bound"),
    ("shape np>=0\n", "expected 'bucket'"),
    ("".join(f"bucket np>={i}\n" for i in range(17)), "more than 16 buckets"),
])
def test_malformed_tables_are_rejected_with_their_line(tunebin, text, msg):
    rc, out = run(tunebin, text)
    assert rc == 3 and out.startswith("ERR tuning stdin line ") and msg in out, out


def test_committed_table_parses(tunebin):
    """the table shipped beside the library is well-formed (it is read at every context creation)"""
    import glob

    files = glob.glob(os.path.join(ROOT, "postgis-cuda_amd", "tuning", "*.txt"))
    if not files:
        pytest.skip("no tuning table committed")
    for f in files:
        rc, out = run(tunebin, open(f).read())
        assert rc == 0 and out.splitlines()[1].startswith("bucket np>=0 "), (f, out)
