"""The ISO_C_BINDING interface module (include/mldgpu.f90) compiled with amdflang against a stand-in
for the reference's mld_mpi module, and a Fortran program that fills mld_params from fit24's ACE keys
and calls mld_ace_basis_review through it (host only, no device): the review must read back fit24's
894 functions and D = 3576 (REF/ndm.out), which checks the bind(c) type against the C struct."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

STUB = """module mld_mpi
  implicit none
contains
  subroutine mld_mpi_abort(msg)
    character(*), intent(in) :: msg
    print '(a)', trim(msg)
    stop 3
  end subroutine
end module
"""

PROGRAM = """program review_fit24
  use iso_c_binding
  use mldgpu_c
  implicit none
  type(mld_params) :: p
  integer(c_int) :: review(42), ace_dim, d, rc
  p%descriptor_type = MLD_DESC_ACE
  p%j_max = 0d0; p%r_cut = 5d0; p%inv_r0_input = 0d0
  p%lbso4_diag = 0; p%nchannel = 1; p%nspecies = 4
  p%weight = c_null_ptr; p%weight_3ch = c_null_ptr; p%r_cut_pair_in = c_null_ptr
  p%r_cut_width_in = 0.4d0; p%precision = MLD_PREC_FP64; p%device = 0
  p%ace_numax = 6
  p%ace_active = [1, 1, 1, 1, 1, 0]
  p%ace_nmax = [14, 2, 1, 1, 1, 1]
  p%ace_lmax = [0, 2, 1, 1, 1, 1]
  p%ace_lambda = 3d0; p%ace_radial_poly = 2; p%ace_gencg = 2; p%ace_chem = 1
  p%ace_npoints_spline = 1000; p%r_cut_width = 0.5d0; p%ace_r_cut_in = 1.7d0
  rc = mld_ace_basis_review(p, review, ace_dim, d)
  print '(3(i0,1x))', rc, ace_dim, d
  print '(7(i0,1x))', review(29:35)
  p%ace_chem = 2
  rc = mld_ace_basis_review(p, review, ace_dim, d)
  call mld_check(c_null_ptr, rc, 'review')
end program
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")
def test_fortran_module_compiles_and_binds_mld_params(tmp_path):
    lib = os.path.join(ROOT, "mld_amd", "libmldgpu.so")
    assert os.path.exists(lib), "build the library first (python -m mld_amd.build)"
    (tmp_path / "mld_mpi.f90").write_text(STUB)
    (tmp_path / "prog.f90").write_text(PROGRAM)
    src = os.path.join(ROOT, "include", "mldgpu.f90")

    def run(cmd):
        r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (cmd, r.stdout, r.stderr)

    run([FLANG, "-c", "mld_mpi.f90"])
    run([FLANG, "-c", src, "-o", "mldgpu.o"])
    run([FLANG, "prog.f90", "mldgpu.o", "mld_mpi.o", "-L", os.path.dirname(lib), "-lmldgpu",
         f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", "prog"])
    r = subprocess.run([str(tmp_path / "prog")], capture_output=True, text=True, timeout=120)
    lines = r.stdout.split("\n")
    assert lines[0].split() == ["0", "894", "3576"], r.stdout
    assert lines[1].split() == ["1024", "1", "3", "676", "476", "396", "476"], r.stdout  # fit24 nu = 5
    # ace_chem = 2 (ACE_CHEM_TS) is refused; mld_check aborts with the library's message
    assert r.returncode == 3 and "ace_chem" in lines[2], (r.returncode, r.stdout)
