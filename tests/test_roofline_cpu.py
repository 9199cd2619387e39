"""The per-kernel byte model and table tool (no GPU)."""
import json
import sqlite3

from polymps_amd import roofline as R


def test_kernel_names():
    assert R._key("void mpsgpu::kPpeAssemble<true, true, true>(mpsgpu::KArgs, double*)") == \
        "kPpeAssemble<true, true, true>"
    assert R._key("mpsgpu::kMxSpmv(mpsgpu::KArgs, mpsgpu::CgPtrs)") == "kMxSpmv"
    assert R._key("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_spmv_bytes_match_the_library_formula():
    """The SpMV figures are the ones the library's profiling counts per
    launch with the LDS windows (spmvBytes in ppe.hip): FP32 5 nnz + 2 nlc +
    16 n, FP64 9 nnz + 2 nlc + 28 n (one-byte window codes, two-byte window
    escapes)."""
    s = {"n": 1000, "L": 120000, "LS": 30000, "nnz": 117000, "nesc": 900, "nlc": 9500}
    m = R.model(s)
    assert m["kMxSpmv"][0] == 5 * 117000 + 2 * 9500 + 16 * 1000
    assert m["kCgSpmv"][0] == 9 * 117000 + 2 * 9500 + 28 * 1000
    # every list consumer counts its list at 4 B per entry
    assert m["kViscosity<true>"][0] == 4 * 120000 + 140 * 1000
    assert m["kCorrectionMatrix<true>"][0] == 4 * 30000 + 136 * 1000


def test_working_dispatches_drop_gated_launches():
    d = [300_000] * 10 + [3_000] * 5
    assert R._working(d) == [300_000] * 10


def test_table_from_databases(tmp_path):
    """A table from minimal databases of the rocprofv3 views it reads."""
    tr = tmp_path / "t.db"
    c = sqlite3.connect(tr)
    c.execute("create table kernels (name text, duration integer)")
    c.executemany("insert into kernels values (?, ?)",
                  [("mpsgpu::kMxSpmv(mpsgpu::KArgs, mpsgpu::CgPtrs)", 250_000)] * 4)
    c.commit()
    pm = {}
    for cn, v in (("FETCH_SIZE", 500.0), ("WRITE_SIZE", 10.0)):
        p = tmp_path / f"{cn}.db"
        c = sqlite3.connect(p)
        c.execute("create table counters_collection (kernel_name text, duration integer, value real, "
                  "counter_name text)")
        c.executemany("insert into counters_collection values (?, ?, ?, ?)",
                      [("mpsgpu::kMxSpmv(mpsgpu::KArgs, mpsgpu::CgPtrs)", 250_000, v, cn),
                       ("__amd_rocclr_copyBuffer", 100_000, 1024.0 * (0.5 if cn == "FETCH_SIZE" else 1.0), cn)])
        c.commit()
        pm[cn] = p
    log = tmp_path / "line.log"
    log.write_text("noise\n" + json.dumps({"metric": "x", "sizes": {
        "particles": 1000, "list_entries": 120000, "list_entries_small": 30000,
        "nnz": 117000, "nnz_escaped": 900}}) + "\n")

    class A:
        trace, fetch, write, sq = str(tr), str(pm["FETCH_SIZE"]), str(pm["WRITE_SIZE"]), None
        copy_fetch, copy_write, copy_bytes, sizes, top = str(pm["FETCH_SIZE"]), str(pm["WRITE_SIZE"]), \
            1024.0 * 1024.0, str(log), 5
    out = R.table(A)
    row = [l for l in out.splitlines() if l.startswith("| kMxSpmv")][0].split("|")
    alg = 6 * 117000 + 4 * 900 + 16 * 1000
    assert row[6].strip() == f"{alg / 1e6:.0f}"
    # FETCH x 2 and WRITE x 1 from the copy: (500 * 2 + 10) KiB
    assert row[9].strip() == f"{(500 * 2 + 10) * 1024 / 1e6:.0f}"
