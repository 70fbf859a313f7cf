"""The bench's readers of the committed rocprofv3 summaries (CPU only): the
traffic figure and the SQ counters the rooflines take from them."""
import os

import bench


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
