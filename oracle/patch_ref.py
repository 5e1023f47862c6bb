#!/usr/bin/env python3
"""Scratch-copy edits for building the reference 3DMHD-CfCA/MCF3DSGincfca3.f90.

Test infrastructure: writes an edited COPY of the reference source into a
temporary build directory (deleted by build_ref.sh after compiling); the
reference tree itself is never modified and no reference source is kept in
this repository.

Edits (each recorded in the fixture manifest written by ref_runner.py):
  1. block size: the compile-time ndx/ndy/ndz/ndmax (M3:3) is set to NDX, so
     that N/NSPLT <= NDX-2 holds for the grid being run;
  2. run directory: the hard-coded '/work/maedarn/3DMHD/samplecnv2/' prefix
     (M3:134,302,309,...) becomes the process working directory;
  3. cooling (not with the "chem" option): the chemistry-network Fcool body
     (M3:2825-2881) is replaced by the Koyama-Inutsuka form of
     1DMHDKIcool/1DMHDwCC3KIcool.f90:1165-1171 with N = U(i,j,k,1)/1.27, the
     same substitution the reference's own
     3DMHD-feedback/MCF3DSGincfca3-feedback.f90:3586 makes; the "chem"
     build keeps the reference's own chemistry Fcool;
  4. (only with the "dtf" option) the jump over the read of the
     inhomogeneous density field (`goto 119` at M3:404) is removed, so INITIA
     reads DTF/D{rank}.dat (M3:405-431) and the x-inflow ghost density is the
     DTF field shifted by -v t (M3:818-826, 854-861).

With the "grav" option the self-gravity file 3DMHD-CfCA/selfG_MPIfull8.f90
is edited too (patch_sg8): mglin's automatic arrays (SG8:155-158) are set to
zero on entry and PB's boundary arrays bphi1/bphi2 right after their
ALLOCATE (SG8:1066, 1080).  The unedited code reads some of them before
writing them (the y/z ghost planes of the coarsest split level, which
`divide` skips on the first ranks, SG8:369-382), so its potential depends on
memory left by earlier calls: over 3 steps at 32^3 on 8 ranks the unedited
build and the zero-filled one differ by 7.8e-10 in Phi (0.15), and filling
the arrays with 1e30 instead gives NaN.

usage: patch_ref.py SRC DST NDX [dtf] [chem]
       patch_ref.py --sg8 SRC DST
"""
import re
import sys

KI_FCOOL = """SUBROUTINE Fcool(CooL,T,i,j,k)
USE comvar
double precision :: CooL,T,N,AbsL,T0,Laml,Lamc
AbsL = 3.94656d1
T0   = 1.0d3
N    = U(i,j,k,1)/1.27d0
Laml = 1.0d7 * dexp( -1.184d5/(T*T0+1.0d3) )
Laml = dmin1(Laml,5.d3)
Lamc = 1.4d-2 * dsqrt(T*T0) * dexp( -9.2d1/(T*T0) )
CooL = ( N**2.d0 * ( Laml + Lamc ) - N ) * AbsL
END SUBROUTINE Fcool
"""


def patch(text: str, ndx: int, dtf: bool = False, chem: bool = False) -> str:
    text, n = re.subn(r"^INTEGER, parameter :: ndx=\d+, ndy=\d+, ndz=\d+, ndmax=\d+, Dim=3.*$",
                      f"INTEGER, parameter :: ndx={ndx}, ndy={ndx}, ndz={ndx}, ndmax={ndx}, Dim=3",
                      text, count=1, flags=re.M)
    assert n == 1, "block-size line not found"
    text = text.replace("'/work/maedarn/3DMHD/samplecnv2/", "'")
    if not chem:
        start = text.index("SUBROUTINE Fcool(CooL,T,i,j,k)")
        end = text.index("END SUBROUTINE Fcool", start) + len("END SUBROUTINE Fcool")
        text = text[:start] + KI_FCOOL.rstrip("\n") + text[end:]
    if dtf:
        text, n = re.subn(r"^(  DTF\(:,:,:\) = dinit1\n)  goto 119\n", r"\1", text, count=1, flags=re.M)
        assert n == 1, "DTF read skip not found"
    return text


def patch_sg8(text: str) -> str:
    """Zero-fill mglin's arrays on entry and PB's bphi1/bphi2 after ALLOCATE."""
    a = "character(3) Nfinal\n\n!Pre-BC for rho is necessary"
    assert text.count(a) == 1, "mglin declarations not found"
    text = text.replace(a, "character(3) Nfinal\n"
                        "cphi1=0.d0; crho1=0.d0; cres1=0.d0; crhs1=0.d0\n"
                        "cphi2=0.d0; crho2=0.d0; cres2=0.d0; crhs2=0.d0\n"
                        "!Pre-BC for rho is necessary")
    for arr in ("bphi1", "bphi2"):
        b = f"ALLOCATE({arr}(Needb,2))\n"
        assert text.count(b) == 1, b
        text = text.replace(b, b + f"{arr}=0.d0\n")
    return text


if __name__ == "__main__":
    if sys.argv[1] == "--sg8":
        with open(sys.argv[2]) as f:
            t = f.read()
        with open(sys.argv[3], "w") as f:
            f.write(patch_sg8(t))
        sys.exit(0)
    src, dst, ndx = sys.argv[1], sys.argv[2], int(sys.argv[3])
    opts = sys.argv[4:]
    assert all(o in ("dtf", "chem", "grav") for o in opts), opts
    with open(src) as f:
        t = f.read()
    with open(dst, "w") as f:
        f.write(patch(t, ndx, "dtf" in opts, "chem" in opts))
