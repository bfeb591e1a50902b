"""The C-ABI surface of the built library (CPU only, no device calls).

* Every function declared in include/*.h is exported by libblis_amd.so.
* Enum values and obj_t layout equal BLIS's (frame/include/bli_type_defs.h:
  254-520 for the bit encodings, :1233-1263 for obj_t), checked by compiling
  a probe against include/blis.h with the system C compiler.
* The netlib driver build needs only blis.h for its integer types, which is
  the drop-in property for BLAS callers.
"""
from __future__ import annotations

import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from tests.conftest import ROOT

LIB = ROOT / "blis_amd" / "lib" / "libblis_amd.so"
HEADERS = [ROOT / "include" / "blis.h", ROOT / "include" / "blis_amd.h",
           ROOT / "include" / "cblas.h"]


def declared_functions():
    """Function names declared by the headers, after the C preprocessor (some
    families are declared through macros, one expansion per datatype)."""
    names = set()
    for h in HEADERS:
        text = subprocess.run(["gcc", "-E", "-P", "-DBLIS_IS_BUILDING_LIBRARY",
                               f"-I{ROOT / 'include'}", "-x", "c", str(h)],
                              capture_output=True, text=True, check=True).stdout
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{[^}]*\}", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(\s*(?:const\s+)?[A-Za-z_]", text):
            name = m.group(1)
            if name.startswith(("bli_", "bli_amd_", "cblas_")) or name.endswith("_"):
                if name in ("defined", "__aligned__"):
                    continue
                names.add(name)
    return names


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_built():
    assert LIB.exists(), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    decl = declared_functions()
    assert {"bli_dgemm", "dgemm_", "bli_dgemm_batch", "dgemm_batch_", "bli_gemm", "bli_dtrsm",
            "dtrsm_", "bli_ddotv", "ddot_", "bli_daxpyv", "daxpy_", "bli_dgemv", "dgemv_",
            "bli_amd_gemm_batch_plan_execute"} <= decl
    missing = sorted(decl - exported())
    assert not missing, missing


def test_constants_exported():
    syms = exported()
    for c in ("BLIS_ONE", "BLIS_ZERO", "BLIS_MINUS_ONE", "BLIS_TWO", "BLIS_MINUS_TWO"):
        assert c in syms


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "blis.h"
int main(void) {
  printf("%d %d %d %d\n", BLIS_NO_TRANSPOSE, BLIS_TRANSPOSE, BLIS_CONJ_NO_TRANSPOSE, BLIS_CONJ_TRANSPOSE);
  printf("%d %d %d %d\n", BLIS_ZEROS, BLIS_LOWER, BLIS_UPPER, BLIS_DENSE);
  printf("%d %d %d %d\n", BLIS_LEFT, BLIS_RIGHT, BLIS_NONUNIT_DIAG, BLIS_UNIT_DIAG);
  printf("%d %d %d %d %d %d\n", BLIS_FLOAT, BLIS_DOUBLE, BLIS_SCOMPLEX, BLIS_DCOMPLEX, BLIS_INT, BLIS_CONSTANT);
  printf("%d %d %d %d\n", BLIS_GENERAL, BLIS_HERMITIAN, BLIS_SYMMETRIC, BLIS_TRIANGULAR);
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(obj_t), offsetof(obj_t, info),
         offsetof(obj_t, elem_size), offsetof(obj_t, buffer), offsetof(obj_t, rs),
         offsetof(obj_t, scalar), offsetof(obj_t, m_padded), offsetof(obj_t, n_panel));
  printf("%zu %zu %zu\n", sizeof(dim_t), sizeof(inc_t), sizeof(f77_int));
  return 0;
}
"""

# Values from BLIS's definitions (bli_type_defs.h:254-520): trans bit 3,
# conj bit 4, uplo bits 5-7, unit-diag bit 8, struc bits 20-21.
EXPECTED = [
    "0 8 16 24",
    "0 192 96 224",
    "0 1 0 256",
    "0 2 1 3 4 5",
    "0 1048576 2097152 3145728",
    # obj_t on LP64: root 8, off 16, dim 16, diag_off 8 -> info at 48;
    # info2 4 -> elem_size 56; buffer 64; rs 72; cs 80; is 88; scalar 96;
    # m_padded 112 ... n_panel 152; size 160.
    "160 48 56 64 72 96 112 152",
    "8 8 4",
]


def test_enum_and_layout_match_blis():
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "probe.c"
        exe = Path(td) / "probe"
        src.write_text(PROBE)
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                       check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.strip().splitlines() == EXPECTED


def test_header_compiles_as_cxx_and_hides_blas_for_netlib():
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "probe.cc"
        src.write_text('#include "blis.h"\n#include "blis_amd.h"\nint main(){return 0;}\n')
        subprocess.run(["g++", "-std=c++11", f"-I{ROOT / 'include'}", str(src), "-o",
                        str(Path(td) / "p")], check=True)
        # With BLIS_VIA_BLASTEST the BLAS prototypes are hidden, so a driver
        # may declare dgemm_ with its own (f2c) signature.
        src2 = Path(td) / "netlib_like.c"
        src2.write_text('#define BLIS_VIA_BLASTEST\n#include "blis.h"\n'
                        'extern int dgemm_(char*, char*, int*, int*, int*, double*, double*, '
                        'int*, double*, int*, double*, double*, int*, int, int);\n'
                        'int main(void){return 0;}\n')
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", "-c", str(src2), "-o",
                        str(Path(td) / "n.o")], check=True)


def test_python_obj_t_mirror_matches():
    import ctypes
    import blis_amd
    assert ctypes.sizeof(blis_amd.obj_t) == 160
    assert blis_amd.obj_t.info.offset == 48
    assert blis_amd.obj_t.buffer.offset == 64
    assert blis_amd.obj_t.scalar.offset == 96


def test_cblas_declared_and_exported():
    decl = declared_functions()
    want = {"cblas_dgemm", "cblas_sgemm", "cblas_zgemm", "cblas_cgemm", "cblas_dgemm_batch",
            "cblas_zgemm_batch", "cblas_dtrsm", "cblas_dtrmm", "cblas_dsymm", "cblas_dsyrk",
            "cblas_dsyr2k", "cblas_dgemmt", "cblas_dgemv", "cblas_ddot", "cblas_daxpy",
            "cblas_xerbla"}
    assert want <= decl
    assert want <= exported()


@pytest.mark.parametrize("call,param,routine", [
    ("cblas_dgemm(99, CblasNoTrans, CblasNoTrans, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2)",
     1, "cblas_dgemm"),
    ("cblas_dgemm(CblasColMajor, 7, CblasNoTrans, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2)",
     2, "cblas_dgemm"),
    ("cblas_dgemm(CblasColMajor, CblasNoTrans, 7, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2)",
     3, "cblas_dgemm"),
    ("cblas_dtrsm(CblasRowMajor, CblasLeft, 7, CblasNoTrans, CblasUnit, 2, 2, 1.0, a, 2, a, 2)",
     3, "cblas_dtrsm"),
    ("cblas_dsyrk(CblasRowMajor, 9, CblasNoTrans, 2, 2, 1.0, a, 2, 0.0, a, 2)",
     3, "cblas_dsyrk"),
])
def test_cblas_argument_errors(tmp_path, call, param, routine):
    """An illegal enum argument is reported by cblas_xerbla with the
    reference's parameter number and ends the process with exit(-1)
    (frame/compat/cblas/src/cblas_xerbla.c). No device is touched: the
    check comes before any computation."""
    src = tmp_path / "probe.c"
    src.write_text(f"""#include "cblas.h"
int main(void) {{ double a[4] = {{0}}; {call}; return 0; }}
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-o", str(exe), str(src), f"-I{ROOT / 'include'}",
                    f"-L{LIB.parent}", "-lblis_amd", f"-Wl,-rpath,{LIB.parent}"], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 255, (p.returncode, p.stderr)
    assert f"Parameter {param} to routine {routine} was incorrect" in p.stderr, p.stderr


def test_registry_queries(tmp_path):
    """bli_gks_query_cntx and the bli_cntx_* queries from a C program built
    against include/blis.h (no device calls): the registered blocksizes are
    the tile table's (FP64 tile 6: 128 x 128 workgroups of 64 x 32 waves, BK
    16; FP32 tile 12: 256 x 128 of 64 x 64, BK 16), the gemm microkernel entry
    is bli_?gemm_ukernel, and the small-problem thresholds bound the sup
    path (frame/base/bli_cntx.h:186-193)."""
    src = tmp_path / "reg.c"
    src.write_text(r'''#include <stdio.h>
#include "blis.h"
int main(void) {
    const cntx_t* c = bli_gks_query_cntx();
    const num_t dts[4] = { BLIS_FLOAT, BLIS_SCOMPLEX, BLIS_DOUBLE, BLIS_DCOMPLEX };
    const bszid_t ids[6] = { BLIS_KR, BLIS_MR, BLIS_NR, BLIS_MC, BLIS_KC, BLIS_NC };
    for (int d = 0; d < 4; ++d) {
        for (int i = 0; i < 6; ++i)
            printf("%ld ", (long)bli_cntx_get_blksz_def_dt(dts[d], ids[i], c));
        printf("| %ld %ld\n", (long)bli_cntx_get_blksz_max_dt(dts[d], BLIS_MC, c),
               (long)bli_cntx_get_bmult_dt(dts[d], BLIS_MC, c));
    }
    printf("ukr %d %d\n", bli_cntx_get_ukr_dt(BLIS_DOUBLE, BLIS_GEMM_UKR, c) == (void_fp)bli_dgemm_ukernel,
           bli_cntx_get_ukr_dt(BLIS_FLOAT, BLIS_GEMM_UKR, c) == (void_fp)bli_sgemm_ukernel);
    printf("rows %d\n", (int)bli_cntx_ukr_prefers_rows_dt(BLIS_DOUBLE, BLIS_GEMM_UKR, c));
    printf("sup %d %d\n", (int)bli_cntx_l3_sup_thresh_is_met(BLIS_DOUBLE, 64, 64, 64, c),
           (int)bli_cntx_l3_sup_thresh_is_met(BLIS_DOUBLE, 4096, 4096, 4096, c));
    printf("impl %s %s %s %s\n", bli_info_get_gemm_impl_string(BLIS_DOUBLE),
           bli_info_get_gemm_impl_string(BLIS_DCOMPLEX),
           bli_info_get_gemm_ukr_impl_string(BLIS_NAT, BLIS_DOUBLE),
           bli_arch_string(bli_arch_query_id()));
    printf("ind %d\n", bli_gks_query_ind_cntx(BLIS_1M) == c);
    printf("info %ld %ld %ld\n", (long)bli_info_get_int_type_size(),
           (long)bli_info_get_blas_int_type_size(), (long)bli_info_get_enable_cblas());
    return 0;
}
''')
    exe = tmp_path / "reg"
    subprocess.run(["gcc", "-o", str(exe), str(src), f"-I{ROOT / 'include'}",
                    f"-L{LIB.parent}", "-lblis_amd", f"-Wl,-rpath,{LIB.parent}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True,
                         timeout=60).stdout.splitlines()
    assert out[0] == "4 64 64 256 16 128 | 256 64"      # float: tile 12
    assert out[2] == "4 64 32 128 16 128 | 128 64"      # double: tile 6
    assert out[1].startswith("4 32 64 128 8 128")      # scomplex: 1m on the float tile
    assert out[3].startswith("4 32 32 64 8 128")       # dcomplex: 1m on the double tile
    assert out[4] == "ukr 1 1"
    assert out[5] == "rows 0"
    assert out[6] == "sup 1 0"
    assert out[7] == "impl native 1m optimzd gfx950"
    assert out[8] == "ind 1"
    assert out[9] == "info 64 32 1"
