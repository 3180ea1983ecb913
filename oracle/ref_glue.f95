! Test infrastructure only: C entry points (bind(C), explicit sizes) over
! the reference's own helper routines, which take assumed-shape arrays and
! optional arguments and so cannot be called from C directly.  The routines
! themselves are compiled from their sources where they lie in the
! reference tree (oracle/Makefile, target `ref`); nothing here restates them.
! The transforms are not among them: they call FFTW, which the image lacks.

subroutine ref_dhaj(n, extend, aj, status) bind(C, name="ref_dhaj")
    ! src/DHaj.f95: aj(1 : n + extend)
    use iso_c_binding
    use SHTOOLS, only: DHaj
    implicit none
    integer(c_int), value :: n, extend
    real(c_double), intent(out) :: aj(*)
    integer(c_int), intent(out) :: status
    call DHaj(n, aj(1:n + extend), extend=extend, exitstatus=status)
end subroutine ref_dhaj

subroutine ref_preglq(x1, x2, n, zero, w, status) bind(C, name="ref_preglq")
    ! src/PreGLQ.f95: n nodes and weights on [x1, x2]
    use iso_c_binding
    use SHTOOLS, only: PreGLQ
    implicit none
    real(c_double), value :: x1, x2
    integer(c_int), value :: n
    real(c_double), intent(out) :: zero(*), w(*)
    integer(c_int), intent(out) :: status
    call PreGLQ(x1, x2, n, zero(1:n), w(1:n), exitstatus=status)
end subroutine ref_preglq

subroutine ref_shglq(lmax, zero, w, status) bind(C, name="ref_shglq")
    ! src/SHGLQ.f95 without the plx table: lmax + 1 nodes and weights
    use iso_c_binding
    use SHTOOLS, only: SHGLQ
    implicit none
    integer(c_int), value :: lmax
    real(c_double), intent(out) :: zero(*), w(*)
    integer(c_int), intent(out) :: status
    call SHGLQ(lmax, zero(1:lmax + 1), w(1:lmax + 1), exitstatus=status)
end subroutine ref_shglq

subroutine ref_shglq_plx(lmax, zero, w, plx, norm, csphase, cnorm, status) bind(C, name="ref_shglq_plx")
    ! src/SHGLQ.f95 with the plx table: plx(lmax + 1, (lmax + 1)(lmax + 2)/2)
    ! in Fortran order
    use, intrinsic :: iso_c_binding
    use SHTOOLS, only: SHGLQ
    implicit none
    integer(c_int), value :: lmax, norm, csphase, cnorm
    real(c_double), intent(out) :: zero(*), w(*)
    real(c_double), intent(out) :: plx(lmax + 1, *)
    integer(c_int), intent(out) :: status
    call SHGLQ(lmax, zero(1:lmax + 1), w(1:lmax + 1), plx=plx(:, 1:(lmax + 1) * (lmax + 2) / 2), norm=norm, &
               csphase=csphase, cnorm=cnorm, exitstatus=status)
end subroutine ref_shglq_plx

subroutine ref_glqgridcoord(lmax, extend, lat, lon, nlat, nlong, status) bind(C, name="ref_glqgridcoord")
    ! src/GLQGridCoord.f95: lat(1 : lmax + 1), lon(1 : 2 lmax + 1 + extend)
    use iso_c_binding
    use SHTOOLS, only: GLQGridCoord
    implicit none
    integer(c_int), value :: lmax, extend
    real(c_double), intent(out) :: lat(*), lon(*)
    integer(c_int), intent(out) :: nlat, nlong, status
    call GLQGridCoord(lat(1:lmax + 1), lon(1:2 * lmax + 1 + extend), lmax, nlat, nlong, extend=extend, &
                      exitstatus=status)
end subroutine ref_glqgridcoord

subroutine ref_plm(kind, lmax, z, csphase, cnorm, p, status) bind(C, name="ref_plm")
    ! kind 1: src/PlmBar.f95, 2: src/PlmSchmidt.f95, 3: src/PlmON.f95 (with
    ! cnorm), 4: src/PLegendreA.f95 (no cnorm); p(1 : (lmax+1)(lmax+2)/2)
    ! in the packed order l(l+1)/2 + m + 1
    use iso_c_binding
    use SHTOOLS, only: PlmBar, PlmSchmidt, PlmON, PLegendreA
    implicit none
    integer(c_int), value :: kind, lmax, csphase, cnorm
    real(c_double), value :: z
    real(c_double), intent(out) :: p(*)
    integer(c_int), intent(out) :: status
    integer :: np
    np = (lmax + 1) * (lmax + 2) / 2
    select case (kind)
    case (1)
        call PlmBar(p(1:np), lmax, z, csphase=csphase, cnorm=cnorm, exitstatus=status)
    case (2)
        call PlmSchmidt(p(1:np), lmax, z, csphase=csphase, cnorm=cnorm, exitstatus=status)
    case (3)
        call PlmON(p(1:np), lmax, z, csphase=csphase, cnorm=cnorm, exitstatus=status)
    case (4)
        call PLegendreA(p(1:np), lmax, z, csphase=csphase, exitstatus=status)
    case default
        status = 2
    end select
end subroutine ref_plm

function ref_normalgravity(lat, gm, omega, a, b) bind(C, name="ref_normalgravity")
    ! src/NormalGravity.f95
    use iso_c_binding
    use SHTOOLS, only: NormalGravity
    implicit none
    real(c_double), value :: lat, gm, omega, a, b
    real(c_double) :: ref_normalgravity
    ref_normalgravity = NormalGravity(lat, gm, omega, a, b)
end function ref_normalgravity

function ref_nglq(degree) bind(C, name="ref_nglq")
    ! src/PreGLQ.f95: NGLQ, NGLQSH, NGLQSHN (degree >= 0: the reference stops otherwise)
    use iso_c_binding
    use SHTOOLS, only: NGLQ
    implicit none
    integer(c_int), value :: degree
    integer(c_int) :: ref_nglq
    ref_nglq = NGLQ(degree)
end function ref_nglq

function ref_nglqsh(degree) bind(C, name="ref_nglqsh")
    use iso_c_binding
    use SHTOOLS, only: NGLQSH
    implicit none
    integer(c_int), value :: degree
    integer(c_int) :: ref_nglqsh
    ref_nglqsh = NGLQSH(degree)
end function ref_nglqsh

function ref_nglqshn(degree, n) bind(C, name="ref_nglqshn")
    use iso_c_binding
    use SHTOOLS, only: NGLQSHN
    implicit none
    integer(c_int), value :: degree, n
    integer(c_int) :: ref_nglqshn
    ref_nglqshn = NGLQSHN(degree, n)
end function ref_nglqshn
