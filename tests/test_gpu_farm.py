"""GPU: the job-table farm (mdkh_run_table) docks every job of a table exactly as the
single-pair program does (byte-identical .out, .detail and .csv files), with one context per
FFT size reused across jobs (ligands of fewer atoms than the context was made for; a job of
another size set up and queued while the previous one runs), retries a failing job before it
reports it, and keeps its host memory bounded on a long table."""
import os
import re
import subprocess
import sys

import pytest

import megadock_amd as M
from megadock_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rows(log):
    return {l.split("\t")[0]: l.split("\t") for l in log.read_text().splitlines() if re.match(r"^\d{5}\t", l)}


def test_farm_matches_single_runs(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    wl = tmp_path / "wl"
    wl.mkdir()
    # the reference's sample table (data/SAMPLE.table: every receptor/ligand combination of
    # 1gcq, which repeats FFT sizes with ligands of different sizes) and two c2 pairs
    pairs = [(r, r), (r, l), (l, r), (l, l)] + [W.make_c2_pair(a, b, str(wl)) for a, b in [(1, 0), (4, 5)]]
    lines = ["%s\t%s\t%s" % (a, b, tmp_path / ("farm_%d" % k)) for k, (a, b) in enumerate(pairs)]
    tb = tmp_path / "jobs.table"
    tb.write_text("TITLE=farm test\nPARAM=-R $1 -L $2 -o $3.out -N 100 -t 2 -O\n" + "\n".join(lines) + "\n")
    log = tmp_path / "master.log"
    print("farm of %d jobs" % len(pairs), flush=True)
    M.run_table(["-tb", str(tb), "-lg", str(log), "-pj", "4"])
    rows = _rows(log)
    assert all(rows["%05d" % (k + 1)][2] == "done" for k in range(len(pairs)))
    for k, (a, b) in enumerate(pairs):
        single = tmp_path / ("single_%d.out" % k)
        M.run(["-R", a, "-L", b, "-o", str(single), "-N", "100", "-t", "2", "-O"])
        for ext in (".out", ".detail", ".csv"):  # the decoys and the detailed output (-O)
            f, g = tmp_path / ("farm_%d%s" % (k, ext)), tmp_path / ("single_%d%s" % (k, ext))
            assert f.read_bytes() == g.read_bytes(), (k, ext)


def test_farm_retries_then_reports_a_failing_job(tmp_path, golden_dir):
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    tb = tmp_path / "jobs.table"
    tb.write_text("PARAM=-R $1 -L $2 -o $3 -N 10\n%s\t%s\t%s\n%s\t%s\t%s\n"
                  % (r, l, tmp_path / "ok.out", r, tmp_path / "missing.pdb", tmp_path / "bad.out"))
    log = tmp_path / "master.log"
    with pytest.raises(M.MdkError, match="1 of 2 jobs failed"):
        M.run_table(["-tb", str(tb), "-lg", str(log), "-rt", "2", "-ot", "3"])
    rows = _rows(log)
    assert rows["00001"][1:3] == ["1", "done"]
    assert rows["00002"][1:3] == ["3", "failed"]  # the first run and two retries
    assert (tmp_path / "ok.out").exists()
    assert len((tmp_path / "ok.out").read_text().splitlines()) == 4 + 10


def test_farm_with_more_fft_sizes_than_live_contexts(tmp_path, golden_dir):
    """Ten FFT sizes, then two of the first ones again, then one size three times in a row,
    with room for about three contexts (-gm 2 GiB): the farm frees the least recently used
    context (never the one whose job is in flight) for each new size, makes a context again
    (with its rotation table) when an evicted size comes back, and lets a job of the
    in-flight job's size wait for it (they share the context).  Every job's .out equals the
    single-pair program's."""
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    sizes = [80, 84, 88, 90, 96, 98, 100, 108, 110, 112, 80, 88, 96, 96, 96]
    lines = ["%s\t%s\t%s\t%d" % (r, l, tmp_path / ("farm_%d.out" % k), nf) for k, nf in enumerate(sizes)]
    tb = tmp_path / "jobs.table"
    tb.write_text("TITLE=sizes\nPARAM=-R $1 -L $2 -o $3 -F $4 -r 24 -N 20 -W 8\n" + "\n".join(lines) + "\n")
    log = tmp_path / "master.log"
    print("farm of %d jobs over %d sizes" % (len(sizes), len(set(sizes))), flush=True)
    M.run_table(["-tb", str(tb), "-lg", str(log), "-pj", "2", "-gm", "2"])
    rows = _rows(log)
    assert all(rows["%05d" % (k + 1)][2] == "done" for k in range(len(sizes)))
    made = re.search(r"contexts made (\d+), freed (\d+)", log.read_text())
    assert made and int(made.group(1)) > 10 and int(made.group(2)) > 0, made
    for k, nf in enumerate(sizes):
        single = tmp_path / ("single_%d.out" % k)
        M.run(["-R", r, "-L", l, "-o", str(single), "-F", str(nf), "-r", "24", "-N", "20", "-W", "8"])
        got = (tmp_path / ("farm_%d.out" % k)).read_bytes()
        assert got == single.read_bytes(), (k, nf)
        assert got.split(b"\t")[0] == str(nf).encode()


def _farm_peak_rss_kib(table, log):
    """Peak resident memory of the farm program run on `table` in a process of its own."""
    dp = os.path.join(os.path.dirname(M.__file__), "bin", "megadock-hip-dp")
    code = ("import resource, subprocess, sys\n"
            "r = subprocess.run(sys.argv[1:], stdout=subprocess.DEVNULL)\n"
            "print(r.returncode, resource.getrusage(resource.RUSAGE_CHILDREN).ru_maxrss)")
    out = subprocess.run([sys.executable, "-c", code, dp, "-tb", str(table), "-lg", str(log), "-pj", "8"],
                         capture_output=True, text=True, timeout=300)
    rc, kib = out.stdout.split()
    assert rc == "0", out.stderr
    return int(kib)


def test_farm_host_memory_stays_bounded_on_a_long_table(tmp_path, golden_dir):
    """A table on which the GPU is the slower stage (the 54000-rotation set: 0.2 s of GPU
    work per job, and 0.6 MB of rotation angles per prepared job; the rotation coefficients
    are made only when a context's table changes): the preparation threads stop at two
    waiting jobs each, so the queue of prepared jobs never holds more than that, and the
    farm's peak memory does not grow with the number of jobs (an unbounded queue would hold
    every job of the longer table)."""
    r = os.path.join(golden_dir, "1gcq_r.pdb")
    l = os.path.join(golden_dir, "1gcq_l.pdb")
    peak = {}
    for n in (20, 96):
        tb = tmp_path / ("long_%d.table" % n)
        tb.write_text("TITLE=long\nPARAM=-R $1 -L $2 -o $3 -D -N 5\n"
                      + "\n".join("%s\t%s\t%s" % (r, l, tmp_path / ("j%d_%d.out" % (n, k))) for k in range(n)) + "\n")
        print("farm of %d jobs of 54000 rotations" % n, flush=True)
        log = tmp_path / ("long_%d.log" % n)
        peak[n] = _farm_peak_rss_kib(tb, log)
        assert len(list(tmp_path.glob("j%d_*.out" % n))) == n
        waiting = re.search(r"at most (\d+) \(limit (\d+)\)", log.read_text())
        assert waiting and int(waiting.group(1)) <= int(waiting.group(2)) == 16, waiting
        if n == 96:
            assert int(waiting.group(1)) == 16  # the preparation ran ahead until the limit held it
    print("peak RSS: %s KiB" % peak, flush=True)
    assert peak[96] - peak[20] < 100 * 1024, peak
