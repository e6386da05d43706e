"""The results harness (ckfft_amd/harness.py): the reference's XML schema and metric on
the CPU; one short run through the library on the GPU."""
import numpy as np
import pytest

from ckfft_amd import harness as H


def test_xml_round_trip_keeps_other_testers(tmp_path):
    path = str(tmp_path / "r.xml")
    root = H.load_results(path)                       # absent file: empty document
    assert root.tag == "ckfft_test" and len(root) == 0
    assert H.record(root, "fft_4096", "ckfft_amd", 0.5, 1e-7) is None
    H.record(root, "fft_4096", "kiss_fft", 0.07, 5.9e-6)
    H.write_results(root, path)
    root2 = H.load_results(path)
    assert H.record(root2, "fft_4096", "ckfft_amd", 0.25, 1e-7) == pytest.approx(0.5)
    el = root2.find("fft_4096")
    assert [c.tag for c in el] == ["ckfft_amd", "kiss_fft"]
    assert float(el.find("kiss_fft").get("time")) == pytest.approx(0.07)


def test_fp64_definitions_and_metric():
    n = 16
    x = np.zeros(2 * n, np.float32)
    x[0] = 1.0                                         # complex delta: flat spectrum
    assert np.allclose(H.fp64_result("fft", x, n)[0::2], 1.0)
    assert np.allclose(H.fp64_result("fft_inv", x, n)[0::2], 1.0)
    r = np.zeros(n, np.float32)
    r[0] = 1.0                                         # real delta: 2 x flat half spectrum
    assert np.allclose(H.fp64_result("fft_real", r, n)[0::2], 2.0)
    spec = H.fp64_result("fft_real", r, n)
    assert np.allclose(H.fp64_result("fft_inv_real", spec, n), 2 * n * r)   # fwd then inv = 2n x
    assert H.rms_error(np.array([1.0, 1.0]), np.array([0.0, 0.0])) == pytest.approx(1.0)


def test_reads_reference_input_layout(tmp_path):
    p = tmp_path / "input.txt"
    p.write_text("-0.5\n0.25\n1.0\n-1.0\n0.75\n")       # odd count: the last value is dropped
    x = H.read_input(str(p))
    assert x.dtype == np.float32 and list(x) == [-0.5, 0.25, 1.0, -1.0]


@pytest.mark.gpu
def test_harness_run_on_gpu(tmp_path):
    import torch
    assert torch.cuda.is_available()
    path = str(tmp_path / "gpu.xml")
    root = H.run(None, path, reps=3, min_n=2048, log=lambda *a: None)
    for n in (4096, 2048):
        for op in ("fft", "fft_inv", "fft_real", "fft_inv_real"):
            el = root.find(f"{op}_{n}")
            assert el is not None and [c.tag for c in el] == list(H.TESTERS)
            for c in el:
                assert float(c.get("time")) > 0
                scale = {"fft": 1, "fft_inv": n, "fft_real": 1, "fft_inv_real": n}[op] * np.sqrt(n)
                assert float(c.get("error")) <= 1e-5 * scale, (op, n, c.tag, c.get("error"))
    assert H.load_results(path).find("fft_real_2048/ckfft_amd_device") is not None
