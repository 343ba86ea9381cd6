"""The draw kernel computes a trail sample's parameter k / ns (render.c:179-185) without a
division per sample: with inv = 1.0 / ns kept per particle, q = k * inv rounded, then one
correction by the exact remainder, r = fma(-q, ns, k) and q' = fma(r, inv, q).  This
checks, for every pair 1 <= k <= ns <= 4096 (ns is clamped to that range), that q' is the
correctly rounded quotient -- the division's own result -- and that the uncorrected
product is not, so the correction is doing work.  The same three operations are compiled
here with gcc (-ffp-contract=off, explicit fma) as the kernel states them."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
int main(void)
{
	long plain = 0, corrected = 0, pairs = 0;
	for (int ns = 1; ns <= 4096; ns++) {
		volatile double inv = 1.0 / ns;
		for (int k = 1; k <= ns; k++) {
			double exact = (double) k / ns;
			double q = (double) k * inv;
			double r = fma(-q, (double) ns, (double) k);
			double q2 = fma(r, inv, q);
			plain += q != exact;
			corrected += q2 != exact;
			pairs++;
		}
	}
	printf("%ld %ld %ld\n", pairs, plain, corrected);
	return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_corrected_ratio_is_the_rounded_quotient(tmp_path):
    c = os.path.join(tmp_path, "ratio.c")
    exe = os.path.join(tmp_path, "ratio")
    with open(c, "w") as f:
        f.write(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, c, "-lm"], check=True)
    pairs, plain, corrected = map(int, subprocess.run([exe], check=True, capture_output=True,
                                                      text=True).stdout.split())
    assert pairs == 4096 * 4097 // 2
    assert plain > 0
    assert corrected == 0
