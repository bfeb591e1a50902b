"""The reference's netlib BLAS test drivers against the GPU library.

blastest/src/{d,s}blat3.c, {d,s}blat2.c and {d,s}blat1.c (compiled from the reference
tree by oracle/Makefile, linked against libblis_amd.so) call dgemm_/dtrsm_/
dgemv_, daxpy_, ddot_ and their s-counterparts with host arrays (the other
level-1 routines of ?blat1 come from the oracle); the library stages them
through device memory. The drivers check every trans/side/uplo/diag case,
alpha/beta special values, edge sizes, untouched padding and the xerbla
error exits (blastest/src/dblat3.c:4109-4317).
"""
from __future__ import annotations

import pytest

from tests.conftest import ROOT, say
from tests.test_oracle import (NETLIB1_CASES, NETLIB_CASES, REF_DIR, check_netlib1_output,
                               check_netlib_output, run_netlib)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("driver,routines,calls", NETLIB_CASES)
def test_netlib_drivers_on_gpu_library(driver, routines, calls):
    exe = REF_DIR / f"{driver}_amd.x"
    if not exe.exists():
        pytest.skip(f"{exe} not built (needs the reference tree at build time)")
    say(f"netlib {driver} on libblis_amd")
    check_netlib_output(run_netlib(driver, "amd"), routines, calls)


@pytest.mark.parametrize("driver,routines", NETLIB1_CASES)
def test_netlib1_drivers_on_gpu_library(driver, routines):
    exe = REF_DIR / f"{driver}_amd.x"
    if not exe.exists():
        pytest.skip(f"{exe} not built (needs the reference tree at build time)")
    say(f"netlib {driver} on libblis_amd")
    check_netlib1_output(run_netlib(driver, "amd"), routines)
