"""The integer-input bandpass kernel divides by its constant norms with a
reciprocal and one fma correction, or with a split reciprocal
(RN(a*rh + RN(a*rl)), two operations), instead of IEEE division.  This checks
both sequences against IEEE a / b on the CPU (same f32 operations, correctly
rounded fma) for the divisors the pipeline uses: every integer dividend up to
2^24 for the boxcar areas, whole binades (every mantissa) and random samples
over 115 binades for the Gaussian norms of several (smoothing_size,
noise_size) pairs.  The library repeats the whole-binade check for each
context's divisors before it uses either sequence."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def div_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("div") / "div_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "div_check.c"), "-lm"], check=True)
    return exe


def _run(exe, b, nrand, ilim, blo, bhi, form="one"):
    out = subprocess.run([exe, repr(float(b)), str(nrand), str(ilim), str(blo), str(bhi), form],
                         capture_output=True, text=True)
    bad, checked = map(int, out.stdout.split())
    return bad, checked


@pytest.mark.parametrize("s", [3, 5, 7, 9, 11, 13, 15])
def test_box_area_division_exact_for_all_integer_sums(div_check, s):
    # boxcar sums of u16 frames are integers below 2^24
    bad, checked = _run(div_check, np.float32(s * s), 0, (s * s) * 65535, 0, 0)
    assert bad == 0 and checked == (s * s) * 65535 + 1


@pytest.mark.parametrize("s,sigma2", [(9, 1.0), (7, 1.0), (13, 1.0), (9, 2.5), (15, 0.7),
                                      (3, 1.0), (11, 4.0)])
def test_gauss_norm_division(div_check, oracle, s, sigma2):
    _, norm = oracle.gauss_weights(s, sigma2)
    bad, checked = _run(div_check, norm, 2_000_000, 70000, 8, 10)
    assert bad == 0 and checked > 2_000_000


@pytest.mark.parametrize("s", [3, 5, 7, 9, 11, 13, 15])
def test_box_area_split_division_exact_for_all_integer_sums(div_check, s):
    bad, checked = _run(div_check, np.float32(s * s), 0, (s * s) * 65535, 0, 0, "split")
    assert bad == 0 and checked == (s * s) * 65535 + 1


@pytest.mark.parametrize("s,sigma2", [(9, 1.0), (7, 1.0), (13, 1.0), (9, 2.5), (15, 0.7),
                                      (3, 1.0), (11, 4.0)])
def test_gauss_norm_split_division(div_check, oracle, s, sigma2):
    _, norm = oracle.gauss_weights(s, sigma2)
    bad, checked = _run(div_check, norm, 2_000_000, 70000, 8, 10, "split")
    assert bad == 0 and checked > 2_000_000


def test_split_form_catches_a_wrong_low_part(div_check, tmp_path):
    # a reciprocal split with its low part dropped is the plain product
    # RN(a * RN(1/b)), which is not IEEE division for b = 81: the check sees it
    src = open(os.path.join(HERE, "div_check.c")).read().replace(
        "rl = (float)(1.0 / (double)b - (double)rh);", "rl = 0.0f;")
    c = tmp_path / "bad.c"
    c.write_text(src)
    exe = str(tmp_path / "bad")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, str(c), "-lm"], check=True)
    bad, _ = _run(exe, np.float32(81.0), 0, 81 * 65535, 0, 0, "split")
    assert bad > 0
