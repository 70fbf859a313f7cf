"""The bench's readers of the committed rocprofv3 summaries (CPU only): the
traffic figure and the SQ counters the rooflines take from them; the serving
batch cap the bench uses."""
import os

import bench
from xin26x_amd import serve


def test_profile_readers(tmp_path):
    p = tmp_path / "k.md"
    p.write_text("| quantity | per launch |\n|---|---|\n| SQ_INSTS_VALU | 61,440,000 |\n"
                 "| SQ_WAVES | 3,072 |\n\ntraffic_bytes_per_launch: 1234\n")
    assert bench.profiled_counter(str(p), "SQ_INSTS_VALU") == 61440000
    assert bench.profiled_counter(str(p), "SQ_INSTS_SALU") is None
    assert bench.profiled_traffic(str(p))[0] == 1234
    r = bench.valu_roofline(str(p), 0.1)                    # 61.44 M instructions in 0.1 ms
    assert r["bound"] == "valu" and abs(r["achieved"] - 614.4) < 1e-6 and abs(r["frac"] - 1.0) < 1e-9
    assert bench.valu_roofline(str(tmp_path / "missing.md"), 0.1) is None


def test_committed_profiles_carry_the_figures_the_bench_reads():
    for path in (bench.PMC_REFINE, bench.PMC_FDCT, bench.PMC_RECON, bench.PMC_QUANT, bench.PMC_BIPRED):
        assert os.path.exists(path), path
        assert bench.profiled_traffic(path)[0], path
    assert bench.profiled_counter(bench.PMC_REFINE, "SQ_INSTS_VALU") > 0


class _Src:
    def __init__(self, n):
        self.blocks = [None] * n


def test_two_chain_cap_splits_the_streams_in_two():
    # the smallest cap under which the streams, taken in order, fill exactly two batches
    srcs = [_Src(n) for n in (3, 3, 3, 10, 10)]
    cap = serve.two_chain_cap(srcs)
    assert cap == 19                                    # (3 + 3 + 3 + 10) | 10 beats 9 | 20
    sizes = [len(s.blocks) for s in srcs]
    batches, cur = [], 0
    for n in sizes:                                     # the server's greedy fill in stream order
        if cur and cur + n > cap:
            batches.append(cur)
            cur = 0
        cur += n
    batches.append(cur)
    assert len(batches) == 2 and max(batches) <= cap
    assert serve.two_chain_cap([_Src(7)]) == 7
