"""bench.py's rocprofv3 summary of the timed executor launches, on a
synthetic kernel trace (CPU only): the launches it picks are the K after the
correctness launch and the W warmup launches, in start order, and other
kernels whose names contain 'exec_kernel' are not counted."""
import csv
import importlib
import statistics


def _trace(path, durations, other_every=0):
    rows, t = [], 1000
    for i, d in enumerate(durations):
        if other_every and i % other_every == 0:
            rows.append({"Kernel_Name": "void srt::mfma_exec_kernel<1>(int)", "Start_Timestamp": t,
                         "End_Timestamp": t + 50000})
            t += 50200
        rows.append({"Kernel_Name": "void srt::exec_kernel<4, 4, 2, false, (anonymous namespace)::Dispatch>(int)",
                     "Start_Timestamp": t, "End_Timestamp": t + d})
        t += d + 100
    path.mkdir()
    with open(path / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def test_timed_launches_are_the_k_after_warmup(tmp_path, capsys):
    bench = importlib.import_module("bench")
    durs = [90000] + [70000] * 5 + [10000 + 100 * i for i in range(20)] + [5000] * 30
    _trace(tmp_path / "p", durs, other_every=7)
    out = tmp_path / "s.txt"
    bench.summarize(str(tmp_path / "p"), str(out), "cmd", timed=(5, 20))
    text = out.read_text()
    want = statistics.mean(d / 1e3 for d in durs[6:26])
    assert "dispatches 7..26 of 56" in text
    assert ("mean %.3f us" % want) in text
    assert "min 10.000 us, max 11.900 us" in text


def test_too_few_launches_prints_no_timed_block(tmp_path):
    bench = importlib.import_module("bench")
    _trace(tmp_path / "p", [10000] * 10)
    out = tmp_path / "s.txt"
    bench.summarize(str(tmp_path / "p"), str(out), "cmd", timed=(5, 20))
    assert "timed executor launches" not in out.read_text()
