"""Timing harness that writes the reference's results format.

The reference's test program (src/test/test.cpp:611-747) times each operation for
N = 4096, 2048, ..., 256 on its input file, one synchronous call per repetition, takes the
median of 500 repetitions in milliseconds, prints name / median / min / max / norm /
change-since-last-run, and keeps the results in an XML file (src/test/test.cpp:566-607)
whose schema is

    <ckfft_test>
        <fft_4096> <ckfft error="0" time="0.054864" /> ... </fft_4096>
        <fft_inv_4096> ... <fft_real_4096> ... <fft_inv_real_4096> ...
    </ckfft_test>

(src/test/results_macos_x86_64.xml:1-5).  This module runs the same operations, in the
same order and on the same data layout, through this library and writes the same schema,
so results line up with the reference's published tables:

  * ``ckfft_amd``         the drop-in call on host buffers, exactly what the reference
                          times (copies to and from the GPU included);
  * ``ckfft_amd_device``  the drop-in call on device buffers (data resident on the GPU);
  * ``ckfft_amd_batch``   the batched call over 256 copies of the signal, time per signal.

``error`` is the reference's metric (RMS over real and imaginary parts,
src/test/test.cpp:140-159) against an FP64 evaluation of the same definition: the complex
DFT and its unnormalised inverse, 2 x the half spectrum for the real forward, and the
unnormalised Hermitian inverse for the real inverse.

    python -m ckfft_amd.harness [--input path/to/input.txt] [--results results.xml]
                                [--reps 500] [--min-n 256]

Without ``--input`` the input is 4096 complex values, real and imaginary parts uniform in
[-1, 1) from a fixed seed.  An existing ``--results`` file is read first: its times give
the change column, and its other testers' entries are kept.
"""
from __future__ import annotations

import argparse
import os
import time
import xml.etree.ElementTree as ET

import numpy as np

TESTERS = ("ckfft_amd", "ckfft_amd_device", "ckfft_amd_batch")
BATCH = 256


def read_input(path: str | None) -> np.ndarray:
    """Complex input as interleaved float32 (re, im, re, im, ...).  The reference's input
    file holds whitespace-separated values read pairwise as (real, imag)
    (src/test/test.cpp:90-107)."""
    if path:
        vals = np.array(open(path).read().split(), dtype=np.float32)
        return vals[: len(vals) // 2 * 2]
    rng = np.random.default_rng(0xCF)
    return rng.uniform(-1.0, 1.0, 2 * 4096).astype(np.float32)


def fp64_result(op: str, x: np.ndarray, n: int) -> np.ndarray:
    """FP64 evaluation of one operation, interleaved like the library's output."""
    x = np.asarray(x, np.float64)
    if op == "fft":
        z = np.fft.fft(x[0::2] + 1j * x[1::2])
    elif op == "fft_inv":
        z = np.fft.ifft(x[0::2] + 1j * x[1::2]) * n
    elif op == "fft_real":
        z = 2.0 * np.fft.rfft(x)
    else:                                            # fft_inv_real: real output
        return np.fft.irfft(x[0::2] + 1j * x[1::2], n) * n
    out = np.empty(2 * len(z))
    out[0::2], out[1::2] = z.real, z.imag
    return out


def rms_error(got: np.ndarray, want: np.ndarray) -> float:
    """src/test/test.cpp:140-159: sqrt(sum of squared re and im differences / (2 count))."""
    d = np.asarray(got, np.float64) - np.asarray(want, np.float64)
    return float(np.sqrt(np.sum(d * d) / len(d)))


def median_ms(fn, reps: int) -> tuple[float, float, float]:
    samples = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        samples.append((time.perf_counter() - t0) * 1e3)
    s = np.sort(samples)
    return float(np.median(s)), float(s[0]), float(s[-1])


def load_results(path: str | None) -> ET.Element:
    if path and os.path.exists(path):
        root = ET.parse(path).getroot()
        if root.tag == "ckfft_test":
            return root
    return ET.Element("ckfft_test")


def record(root: ET.Element, test: str, tester: str, t_ms: float, error: float) -> float | None:
    """Store one result; returns the previous time of this (test, tester), if any."""
    test_el = root.find(test)
    if test_el is None:
        test_el = ET.SubElement(root, test)
    el = test_el.find(tester)
    prev = None
    if el is None:
        el = ET.SubElement(test_el, tester)
    elif el.get("time") is not None:
        prev = float(el.get("time"))
    el.set("error", f"{error:g}")
    el.set("time", f"{t_ms:g}")
    return prev


def write_results(root: ET.Element, path: str) -> None:
    ET.indent(root, space="    ")
    ET.ElementTree(root).write(path, encoding="unicode", xml_declaration=False)


def run(input_path=None, results_path=None, reps: int = 500, min_n: int = 256, max_n: int = 4096,
        testers=TESTERS, log=print) -> ET.Element:
    import torch

    import ckfft_amd as C

    x_all = read_input(input_path)
    root = load_results(results_path)
    dev = torch.device("cuda")
    log(f"\n{reps} iterations")
    n = min(max_n, len(x_all) // 2)
    while n >= min_n:
        x = x_all[: 2 * n].copy()
        ops = [("fft", False, False), ("fft_inv", True, False), ("fft_real", False, True),
               ("fft_inv_real", True, True)]
        results = {}
        for op, inverse, real in ops:
            # inputs follow the reference's chaining: inverses take the forward outputs
            if op == "fft":
                inp = x
            elif op == "fft_inv":
                inp = results["fft"]
            elif op == "fft_real":
                inp = x[:n]
            else:
                inp = results["fft_real"]
            out_len = n if op == "fft_inv_real" else (2 * (n // 2 + 1) if real else 2 * n)
            direction = C.kCkFftDirection_Inverse if inverse else C.kCkFftDirection_Forward
            test = f"{op}_{n}"
            log(f"\n{test}:\n        name    median       min       max      norm    change")
            log("-" * 68)
            base = None
            with C.Context(n, direction) as ctx:
                call = {"fft": ctx.complex_forward, "fft_inv": ctx.complex_inverse,
                        "fft_real": ctx.real_forward}.get(op)
                batch = {"fft": ctx.complex_forward_batch, "fft_inv": ctx.complex_inverse_batch,
                         "fft_real": ctx.real_forward_batch, "fft_inv_real": ctx.real_inverse_batch}[op]
                tmp = np.zeros(2 * (n // 2 + 1), np.float32)
                for tester in testers:
                    if tester == "ckfft_amd":
                        xi = np.ascontiguousarray(inp, np.float32)
                        yo = np.zeros(out_len, np.float32)
                        fn = (lambda: ctx.real_inverse(n, xi, yo, tmp)) if call is None else (lambda: call(n, xi, yo))
                        get = lambda: yo
                    elif tester == "ckfft_amd_device":
                        xi = torch.from_numpy(np.ascontiguousarray(inp, np.float32)).to(dev)
                        yo = torch.zeros(out_len, device=dev)
                        td = torch.zeros(2 * (n // 2 + 1), device=dev)
                        fn = (lambda: ctx.real_inverse(n, xi, yo, td)) if call is None else (lambda: call(n, xi, yo))
                        get = lambda: yo.cpu().numpy()
                    else:
                        xi = torch.from_numpy(np.ascontiguousarray(inp, np.float32)).to(dev).repeat(BATCH, 1)
                        yo = torch.zeros(BATCH, out_len, device=dev)

                        def fn():
                            batch(n, BATCH, xi, yo, None)
                            torch.cuda.synchronize()
                        get = lambda: yo[0].cpu().numpy()
                    assert fn() in (1, None)
                    med, lo, hi = median_ms(fn, reps)
                    if tester == "ckfft_amd_batch":
                        med, lo, hi = med / BATCH, lo / BATCH, hi / BATCH
                    out = np.array(get(), np.float32)
                    if tester == "ckfft_amd":
                        results[op] = out
                    err = rms_error(out, fp64_result(op, inp, n))
                    prev = record(root, test, tester, med, err)
                    base = med if base is None else base
                    change = 100.0 * (med - prev) / prev if prev else 0.0
                    log(f"{tester:>16}: {med:f}  {lo:f}  {hi:f}  {med / base:f}  {change:7.2f}%")
        n //= 2
    if results_path:
        write_results(root, results_path)
        log(f"\nwrote results to {results_path}")
    return root


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--input", help="the reference's src/test/input.txt, or any file of real/imag pairs")
    ap.add_argument("--results", default="results_mi355x.xml", help="XML results file (read, then written)")
    ap.add_argument("--reps", type=int, default=500)
    ap.add_argument("--min-n", type=int, default=256)
    args = ap.parse_args()
    run(args.input, args.results, args.reps, args.min_n)


if __name__ == "__main__":
    main()
