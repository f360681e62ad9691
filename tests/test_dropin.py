"""The reference-named entry points (include/vpic_dropin.h) driven by a
program compiled against lanl/vpic's own headers (oracle/ref/dropin_check.cc,
built into oracle/_ref by the reference build), checked inside that program
against the reference's own kernels: load_interpolator_array, advance_p
(unsorted, after sort_p, with absorbing walls), sort_p, move_p, energy_p,
center_p/uncenter_p, unload_accumulator_array, accumulate_rho_p, the field
kernels through fa->kernel, resident mode and the error handler."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CHECK = os.path.join(ROOT, "oracle", "_ref", "dropin_check")


@pytest.mark.gpu
@pytest.mark.ref
@pytest.mark.skipif(not os.path.exists(REF_CHECK),
                    reason="oracle/_ref (reference build) not present")
def test_dropin_against_reference_kernels():
    print("dropin_check", flush=True)
    out = subprocess.run([REF_CHECK], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout
    assert "0 failed" in out.stdout
