"""empi-hip CLI end to end (needs a GPU): config file in, MP5 book and the
stdout protocol out, checked against the FP64 oracle and the reference's
output conventions (src/empi.cpp:132-164, src/classes.cpp:41-48,83)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from empi_amd import synth
from parity import AMP_RTOL, oracle_following

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "empi_amd", "empi-hip")
TOOL = os.path.join(ROOT, "empi_amd", "empi-host-tool")


def _write_case(tmp, data, cfg_extra, name="sig.bin"):
    data.astype(np.float32).tofile(tmp / name)   # [t][channel] interleaved
    cfg = tmp / "run.cfg"
    cfg.write_text(cfg_extra + f"\nnameOfDataFile {tmp / name}\nnameOfOutputDirectory {tmp}\n")
    return cfg


def _read_book(path):
    out = subprocess.run([TOOL, "readbook", str(path)], capture_output=True, text=True, check=True).stdout
    rows = [line.split() for line in out.splitlines() if not line.startswith(("fs", "epoch"))]
    return [(int(r[0]), np.array([float(v) for v in r[1:]])) for r in rows]


def _book_index(d, params):
    """(k, f, t) of book atoms from their float32 parameters (position and
    scale in samples, frequency in units of Nyquist)."""
    scales = np.array([sc.s * d.fs for sc in d.scales])
    out = []
    for p in params:
        k = int(np.argmin(np.abs(scales - p[3])))
        sc = d.scales[k]
        out.append((k, int(round(p[4] * sc.Nfft / 2)), int(round(p[2] / d.fs * (sc.T - 1) / d.tMax))))
    return np.array(out, dtype=np.int64)


def _book_parity(d, x, variant, iters, percent, params, where):
    """Book atoms (float32 parameters) against the oracle following their
    indices: indices, counts, float32-rounded position, scale, frequency and
    amplitude at every iteration."""
    params = np.asarray(params, np.float64)
    C = x.shape[0]
    n = len(params) // C
    per = params.reshape(C, n, 6)
    index = np.stack([_book_index(d, per[c]) for c in range(C)])
    ora = oracle_following(d, x, variant, iters, percent, index, [n] * C)
    for c in range(C):
        assert int(ora.counts[c]) == n, (where, c, int(ora.counts[c]), n)
        assert np.array_equal(ora.index[c, :n], index[c]), (where, c)
        want = ora.params[c, :n].astype(np.float32).astype(np.float64)
        assert np.allclose(per[c, :, 2:5], want[:, 2:5], rtol=1e-6), (where, c)
        assert np.all(np.abs(per[c, :, 1] / want[:, 1] - 1) < AMP_RTOL), (where, c)
    return n


@pytest.mark.parametrize("per_call", [False, True])
def test_cli_smp_two_epochs(tmp_path, per_call):
    print(f"\n[gpu case] CLI SMP 2 epochs x 2 channels, per_call={per_call}", flush=True)
    N = 1280
    sig = np.stack([np.concatenate([synth.single_channel(seed=10 + c), synth.single_channel(seed=20 + c)])
                    for c in range(2)])                       # [channel][2N]
    cfg = _write_case(tmp_path, sig.T.copy(),
                      "energyError 0.01\nmaximalNumberOfIterations 20\nenergyPercent 99.0\nMP SMP\n"
                      "samplingFrequency 128.0\nnumberOfChannels 2\nselectedChannels 1-2\n"
                      f"numberOfSamplesInEpoch {N}\n")
    args = [CLI] + (["--per-call"] if per_call else []) + [str(cfg)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    assert lines[0] == "START\t1\t2\t20\t99"
    assert lines[-1] == " END"
    assert lines.count("CHANNEL\t0") == 2 and lines.count("CHANNEL\t1") == 2
    book = tmp_path / "sig_smp.b"
    assert book.exists()
    atoms = _read_book(book)
    d = po.Dictionary(128.0, N, 0.01)
    k = 0
    for e in range(2):
        for c in range(2):
            x = sig[c, e * N:(e + 1) * N][None, :].astype(np.float64)
            n = 0
            while k + n < len(atoms) and atoms[k + n][0] == c + 1:
                n += 1
            _book_parity(d, x, po.SMP, 20, 99.0, [p for _, p in atoms[k:k + n]], f"epoch {e} channel {c}")
            k += n
            atom_lines = [ln for ln in lines if ln.startswith("ATOM\t")]
            assert atom_lines[0].split("\t")[:3] == ["ATOM", "0", str(d.entries)]
    assert k == len(atoms)


def test_cli_errors(tmp_path):
    print("\n[gpu case] CLI error codes", flush=True)
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("energyError 2\n")
    r = subprocess.run([CLI, str(cfg)], capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.strip() == "ERROR\tinvalidEnergyErrorValue"
    cfg.write_text("energyError 0.01\n")
    r = subprocess.run([CLI, str(cfg)], capture_output=True, text=True)
    assert r.stdout.strip() == "ERROR\tmissingConfigurationEntry:maximalNumberOfIterations"
    zero = np.zeros((512, 1), np.float32)
    cfg2 = _write_case(tmp_path, zero, "energyError 0.01\nmaximalNumberOfIterations 5\nenergyPercent 99\nMP smp\n"
                                       "samplingFrequency 64\nnumberOfChannels 1\nselectedChannels 1\n", "z.bin")
    r = subprocess.run([CLI, str(cfg2)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and r.stdout.splitlines()[-1] == "ERROR\tsignalIsEmpty"


def test_cli_mmp_variants_name_and_run(tmp_path):
    print("\n[gpu case] CLI MMP3 4 channels", flush=True)
    x = synth.shared_multichannel(C=4, N=1280, fs=128.0, atoms=3, seed=9)
    cfg = _write_case(tmp_path, x.T.copy(), "energyError 0.01\nmaximalNumberOfIterations 8\nenergyPercent 99\n"
                                            "MP MMP3\nsamplingFrequency 128\nnumberOfChannels 4\nselectedChannels 1-4\n")
    r = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert (tmp_path / "sig_mmp.b").exists()
    assert "CHANNEL" not in r.stdout
    atoms = _read_book(tmp_path / "sig_mmp.b")
    assert len(atoms) % 4 == 0
    assert [ch for ch, _ in atoms] == [c + 1 for c in range(4) for _ in range(len(atoms) // 4)]
    _book_parity(po.Dictionary(128.0, 1280, 0.01), x.astype(np.float64), po.MMP3, 8, 99.0,
                 [p for _, p in atoms], "MMP3 book")


@pytest.mark.parametrize("per_batch", [2, 3])
def test_cli_streamed_batches_match_one_batch(tmp_path, per_batch):
    """--epochs-per-batch streams the file in device batches (reader thread,
    reused plan, the last batch shorter): stdout and the book are byte-for-byte
    those of the single batch, for SMP and MMP3."""
    print(f"\n[gpu case] CLI 5 epochs, one batch vs --epochs-per-batch={per_batch}", flush=True)
    N = 1024
    for mp, C in [("SMP", 2), ("MMP3", 3)]:
        x = np.concatenate([synth.shared_multichannel(C=C, N=N, fs=256.0, atoms=3, seed=50 + e) for e in range(5)],
                           axis=1)                           # [channel][5N]
        case = tmp_path / f"{mp}_{per_batch}"
        case.mkdir()
        cfg = _write_case(case, x.T.copy(),
                          f"energyError 0.01\nmaximalNumberOfIterations 12\nenergyPercent 99\nMP {mp}\n"
                          f"samplingFrequency 256\nnumberOfChannels {C}\nselectedChannels 1-{C}\n"
                          f"numberOfSamplesInEpoch {N}\n")
        book = case / f"sig_{mp[:3].lower()}.b"
        one = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, timeout=300)
        assert one.returncode == 0, one.stdout[-2000:]
        one_book = book.read_bytes()
        book.unlink()
        streamed = subprocess.run([CLI, f"--epochs-per-batch={per_batch}", str(cfg)], capture_output=True,
                                  text=True, timeout=300)
        assert streamed.returncode == 0, streamed.stdout[-2000:]
        assert streamed.stdout == one.stdout
        assert book.read_bytes() == one_book
        assert one.stdout.count("START\t") == 5


def _book_epochs(path):
    """{book epoch: [(channel, params)]} from readbook."""
    out = subprocess.run([TOOL, "readbook", str(path)], capture_output=True, text=True, check=True).stdout
    epochs, cur = {}, None
    for line in out.splitlines():
        if line.startswith("fs"):
            continue
        if line.startswith("epoch"):
            cur = int(line.split()[1])
            epochs[cur] = []
        else:
            r = line.split()
            epochs[cur].append((int(r[0]), tuple(float(v) for v in r[1:])))
    return epochs


def test_cli_selected_epochs_and_truncated_files(tmp_path):
    """selectedEpochs (src/empi.cpp:87-98, io.cpp:64-79) decomposes the listed
    epochs in the listed order, each exactly as in the all-epochs run, in one
    batch or streamed one epoch per batch with the next epoch staged on the
    device while the current one runs; a partial last epoch, and a selected
    epoch past the end of the file, end the run with fileIsTruncated after
    the epochs before it (io.cpp:48-54, 64-79)."""
    print("\n[gpu case] CLI selectedEpochs 3,1 and truncated files", flush=True)
    N, E, C = 1024, 4, 2
    x = np.concatenate([synth.shared_multichannel(C=C, N=N, fs=256.0, atoms=3, seed=90 + e) for e in range(E)],
                       axis=1)                                 # [channel][E N]
    base = (f"energyError 0.01\nmaximalNumberOfIterations 10\nenergyPercent 99\nMP SMP\nsamplingFrequency 256\n"
            f"numberOfChannels {C}\nselectedChannels 1-{C}\nnumberOfSamplesInEpoch {N}\n")
    all_dir = tmp_path / "all"
    all_dir.mkdir()
    cfg = _write_case(all_dir, x.T.copy(), base)
    r = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    want = _book_epochs(all_dir / "sig_smp.b")
    assert sorted(want) == [1, 2, 3, 4]
    for extra in ([], ["--epochs-per-batch=1"]):
        sel = tmp_path / f"sel{len(extra)}"
        sel.mkdir()
        cfg = _write_case(sel, x.T.copy(), base + "selectedEpochs 3,1\n")
        r = subprocess.run([CLI] + extra + [str(cfg)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:]
        assert r.stdout.count("START\t") == 2
        got = _book_epochs(sel / "sig_smp.b")
        assert got == {1: want[3], 2: want[1]}, extra
    # a partial last epoch: three full epochs decomposed, then the error
    cut = tmp_path / "cut"
    cut.mkdir()
    cfg = _write_case(cut, x[:, :3 * N + 100].T.copy(), base)
    for extra in ([], ["--epochs-per-batch=1"]):
        r = subprocess.run([CLI] + extra + [str(cfg)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 1 and r.stdout.splitlines()[-1] == "ERROR\tfileIsTruncated", r.stdout[-500:]
        assert r.stdout.count("START\t") == 3, extra
    # a selected epoch past the end
    cfg = _write_case(cut, x.T.copy(), base + "selectedEpochs 2,5\n")
    r = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and r.stdout.splitlines()[-1] == "ERROR\tfileIsTruncated", r.stdout[-500:]
    assert r.stdout.count("START\t") == 1
