"""bench.py's counter arithmetic on the CPU: counter_passes runs its rocprofv3 passes against a stand-in profiler
that writes a results database of rocprofv3's schema (table counters_collection) with chosen values, so the traffic
(FETCH_SIZE and WRITE_SIZE scaled by the 512 MiB copy calibration) and the valu record (VALUBusy, the clock from
GRBM_GUI_ACTIVE over the dispatch's duration, the VALU issue fraction) are checked against figures worked out by
hand here."""
import os
import stat
import sys

import pytest

import bench

# One eager pass as the stand-in reports it: dispatch -> (kernel, start ns, end ns, {counter: value}).
PASS = {
    1: ("void rcg::k_encode<true>(rcg::EncJob const*, unsigned int const*, int)", 0, 4_900_000,
        {"FETCH_SIZE": 300.0, "WRITE_SIZE": 100.0, "VALUBusy": 108.6, "SQ_INSTS_VALU": 2.7e9, "GRBM_GUI_ACTIVE": 7.8e7}),
    2: ("void rcg::k_ssim_ref<256, 0u>(rcg::RefJob const*)", 0, 6_810_000,
        {"FETCH_SIZE": 50.0, "WRITE_SIZE": 70.0, "VALUBusy": 80.4, "SQ_INSTS_VALU": 2.4e9, "GRBM_GUI_ACTIVE": 9.5e7}),
    3: ("void rcg::k_ssim_scale<float, 256, 3u, 1u>(rcg::SsimJob const*)", 0, 55_280_000,
        {"FETCH_SIZE": 1000.0, "WRITE_SIZE": 500.0, "VALUBusy": 91.7, "SQ_INSTS_VALU": 29_248_597_800.0,
         "GRBM_GUI_ACTIVE": 1_017_526_455.0}),
    4: ("void rcg::k_ssim_scale<float, 128, 3u, 2u>(rcg::SsimJob const*)", 0, 2_770_000,
        {"FETCH_SIZE": 200.0, "WRITE_SIZE": 100.0, "VALUBusy": 79.0, "SQ_INSTS_VALU": 1.26e9, "GRBM_GUI_ACTIVE": 4.9e7}),
}
# The copy calibration: 512 MiB copied, read as half its bytes by FETCH_SIZE and all of them by WRITE_SIZE.
COPY = {1: ("__amd_rocclr_copyBuffer", 0, 100_000, {"FETCH_SIZE": 262144.0, "WRITE_SIZE": 524288.0})}

FAKE = r'''#!{python}
import os, sqlite3, sys, json
argv = sys.argv[1:]
pmc = argv[argv.index("--pmc") + 1:argv.index("-d")]
d, name = argv[argv.index("-d") + 1], argv[argv.index("-o") + 1]
child = argv[argv.index("--") + 1:]
table = json.load(open({table!r}))["copy" if "--copy-calibration" in child else "pass"]
os.makedirs(os.path.join(d, "host"), exist_ok=True)
con = sqlite3.connect(os.path.join(d, "host", name + "_results.db"))
con.execute('create table counters_collection (dispatch_id, kernel_name, counter_name, value, start, "end")')
for did, (k, s, e, vals) in table.items():
    for c in pmc:
        # each counter as two instances (e.g. two dies) whose values add up
        con.execute("insert into counters_collection values (?,?,?,?,?,?)", (int(did), k, c, vals[c] / 2, s, e))
        con.execute("insert into counters_collection values (?,?,?,?,?,?)", (int(did), k, c, vals[c] / 2, s, e))
con.commit()
'''


@pytest.fixture
def fake_profiler(tmp_path, monkeypatch):
    import json
    table = tmp_path / "table.json"
    table.write_text(json.dumps({"pass": PASS, "copy": COPY}))
    exe = tmp_path / "rocprofv3"
    exe.write_text(FAKE.format(python=sys.executable, table=str(table)))
    exe.chmod(exe.stat().st_mode | stat.S_IXUSR)
    monkeypatch.setenv("PATH", str(tmp_path) + os.pathsep + os.environ.get("PATH", ""))
    return exe


def test_pmc_run_sums_instances_and_keeps_dispatch_order(fake_profiler, tmp_path):
    rows = bench._pmc_run("VALUBusy SQ_INSTS_VALU", ["--batch-pass-child", "{}"], str(tmp_path), "p")
    assert [r[1] for r in rows] == [1, 2, 3, 4]
    k, _, vals, ns = rows[2]
    assert k == "k_ssim_scale<float, 256, 3u, 1u>" and ns == 55_280_000
    assert vals == {"VALUBusy": 91.7, "SQ_INSTS_VALU": 29_248_597_800.0}


def test_counter_passes_traffic_and_issue_rate(fake_profiler):
    traffic, valu = bench.counter_passes(4, {"wide_nt": 256})
    # Calibration factors: 512 MiB / 262144 KiB = 2 for FETCH_SIZE, 512 MiB / 524288 KiB = 1 for WRITE_SIZE.
    assert traffic["fetch_factor"] == 2.0 and traffic["write_factor"] == 1.0
    # The two FP32 scale launches: mean fetch (1000 + 200) / 2 KiB x 2, mean write (500 + 100) / 2 KiB x 1.
    assert traffic["launches"] == 2
    assert traffic["fetch_bytes"] == pytest.approx(600 * 2 * 1024)
    assert traffic["write_bytes"] == pytest.approx(300 * 1024)
    assert traffic["bytes_per_launch"] == pytest.approx((600 * 2 + 300) * 1024)
    assert traffic["scale0_bytes"] == pytest.approx((1000 * 2 + 500) * 1024)
    assert valu["k_ssim_scale_scale0_pct"] == 91.7 and valu["k_encode_pct"] == 108.6
    assert valu["k_ssim_ref_scale0_pct"] == 80.4
    # 1,017,526,455 busy cycles over 8 dies in 55.28 ms: 2.301 GHz; 29,248,597,800 x 4 / (1024 x 127,190,807) = 0.898.
    assert valu["k_ssim_scale_scale0_clock_ghz"] == pytest.approx(2.301, abs=1e-3)
    assert valu["k_ssim_scale_scale0_issue_frac"] == pytest.approx(0.898, abs=1e-3)


def test_a_failed_pass_gives_no_figures(tmp_path, monkeypatch):
    exe = tmp_path / "rocprofv3"
    exe.write_text(f"#!{sys.executable}\nimport sys\nsys.exit(3)\n")
    exe.chmod(exe.stat().st_mode | stat.S_IXUSR)
    monkeypatch.setenv("PATH", str(tmp_path) + os.pathsep + os.environ.get("PATH", ""))
    assert bench._pmc_run("FETCH_SIZE", ["--copy-calibration", "512"], str(tmp_path), "x") is None
    assert bench.counter_passes(4, {}) == (None, None)
