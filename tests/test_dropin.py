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


@pytest.mark.gpu
@pytest.mark.ref
@pytest.mark.skipif(not os.path.exists(REF_CHECK),
                    reason="oracle/_ref (reference build) not present")
def test_dropin_resident_at_scale():
    """The reference's step through the reference-named entry points in
    resident mode on a 128x64x64 box with 2 x 100 ppc (104,857,600
    particles, density 1): no particle crosses PCIe between steps, so the
    rate is the device's (measured 3.3e10 particle-steps/s; a step that
    uploaded and downloaded the particles would run below 1e9 over
    PCIe)."""
    print("dropin_check --bench", flush=True)
    out = subprocess.run([REF_CHECK, "--bench", "128", "64", "64", "100",
                          "5", "40"], capture_output=True, text=True,
                         timeout=600)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [l for l in out.stdout.splitlines()
            if l.startswith("dropin resident:")][0]
    rate = float(line.split(" particle-steps/s")[0].split()[-1])
    assert rate > 1e10, line
