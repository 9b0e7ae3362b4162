!> iso_c_binding shim between a CHAPSim Fortran host program and the MI355X
!> library libchapsim_gpu.so (C ABI: include/chapsim_gpu.h).
!>
!> Compile it with the same Fortran compiler as the host program, e.g.
!>     amdflang -c -O2 chapsim_gpu.f90
!> and link the program with -L<dir> -lchapsim_gpu.  Every wrapper turns a
!> non-zero status into a message plus ERROR STOP (the reference STOPs on
!> errors, src/WRTHDL.f90:18-29).
!>
!> Drop-in wrappers keep the reference's argument lists:
!>     CALL CHS_TDMAIJJ_nonCYC(A,B,C,R,BCJ,JS,JSZ,IS,ISZ)   (src/TDMAIJJ_nonCYC.f90:67)
!>     CALL CHS_TDMAIJI_nonCYC(A,B,C,R,BCI,IS,ISZ,JS,JSZ)   (src/TDMAIJI_nonCYC.f90:67)
!>     CALL CHS_TDMAIJJ_CYC(A,B,C,R,JS,JSZ,IS,ISZ)          (src/TDMAIJJ_CYC.f90:67)
!>     CALL CHS_TDMAIJI_CYC(A,B,C,R,IS,ISZ,JS,JSZ)          (src/TDMAIJI_CYC.f90:67)
!> and the device-resident substep:
!>     CALL CHS_INIT(...)        once after TDMA_COEF / FFT99_POIS3D_INIT
!>     CALL CHS_PUT_STATE(Q_tg, PR_tg, CONVH0_tg)
!>     CALL CHS_SOLVERRK3_MOM_TG(NS, DT)   in place of SOLVERRK3_MOM_tg(NS)
!>     CALL CHS_GET_STATE(Q_tg, PR_tg, CONVH0_tg)   before output / restart
module chapsim_gpu
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: CHS_OK = 0, CHS_FP64 = 0, CHS_MIXED = 1

    type, bind(c) :: chs_grid_desc
        integer(c_int) :: ncl1, ncl2, ncl3, icase, flowdrvtp, precision
        real(c_double) :: dxi, dzi, dxqi, dzqi, cvisc, volm
        type(c_ptr) :: dyfi, dyci, wff, wfb, amph, acph, apph, amvr, acvr, apvr, ak1, ak3
    end type

    interface
        integer(c_int) function c_create(desc, stream, h) bind(c, name='chs_create')
            import :: c_int, c_ptr, chs_grid_desc
            type(chs_grid_desc), intent(in) :: desc
            type(c_ptr), value :: stream
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function c_destroy(h) bind(c, name='chs_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function c_upload(h, q, pr, convh0) bind(c, name='chs_upload_state')
            import :: c_int, c_ptr
            type(c_ptr), value :: h, q, pr, convh0
        end function
        integer(c_int) function c_download(h, q, pr, convh0) bind(c, name='chs_download_state')
            import :: c_int, c_ptr
            type(c_ptr), value :: h, q, pr, convh0
        end function
        integer(c_int) function c_substep(h, ns, dt) bind(c, name='chs_rk3_substep')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: ns
            real(c_double), value :: dt
        end function
        integer(c_int) function c_divgck(h, v) bind(c, name='chs_divgck')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function c_cfl(h, v) bind(c, name='chs_cfl')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function c_tdma_host(a, b, c, r, bc, n, nlines, es, ls, cyclic) &
                bind(c, name='chs_tdma_batched_host')
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: a, b, c, r, bc
            integer(c_int), value :: n, nlines, cyclic
            integer(c_long_long), value :: es, ls
        end function
        type(c_ptr) function c_last_error() bind(c, name='chs_last_error')
            import :: c_ptr
        end function
    end interface

    type(c_ptr), save :: handle = c_null_ptr

    public :: chs_init, chs_finalize, chs_put_state, chs_get_state, chs_solverrk3_mom_tg, &
              chs_divgck_tg, chs_cfl_tg, chs_tdmaijj_noncyc, chs_tdmaiji_noncyc, &
              chs_tdmaijj_cyc, chs_tdmaiji_cyc

contains

    subroutine chs_check(st, where)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (st == CHS_OK) return
        call c_f_pointer(c_last_error(), msg, [512])
        n = 0
        do while (n < 512)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        end do
        write(*, '(A,A,I3,A,512A1)') where, ': chapsim_gpu error ', st, ': ', msg(1:n)
        error stop 1
    end subroutine

    !> Mesh and coefficient arrays as the reference's modules hold them:
    !> DYFI(1:NCL2), DYCI(1:NND2), YCL2ND_WFF/WFB(1:NND2), AMPH/ACPH/APPH(1:NCL2),
    !> AMVR/ACVR/APVR(1:NND2,3), AK1(1:NCL1), AK3(1:NCL3).
    subroutine chs_init(ncl1, ncl2, ncl3, dxi, dzi, dxqi, dzqi, cvisc, volm, dyfi, dyci, wff, wfb, &
                        amph, acph, apph, amvr, acvr, apvr, ak1, ak3, precision)
        integer, intent(in) :: ncl1, ncl2, ncl3, precision
        real(c_double), intent(in) :: dxi, dzi, dxqi, dzqi, cvisc, volm
        real(c_double), target, intent(in) :: dyfi(:), dyci(:), wff(:), wfb(:), amph(:), acph(:), apph(:)
        real(c_double), target, intent(in) :: amvr(:, :), acvr(:, :), apvr(:, :), ak1(:), ak3(:)
        type(chs_grid_desc) :: d
        d%ncl1 = ncl1; d%ncl2 = ncl2; d%ncl3 = ncl3
        d%icase = 1; d%flowdrvtp = 1; d%precision = precision
        d%dxi = dxi; d%dzi = dzi; d%dxqi = dxqi; d%dzqi = dzqi; d%cvisc = cvisc; d%volm = volm
        d%dyfi = c_loc(dyfi); d%dyci = c_loc(dyci); d%wff = c_loc(wff); d%wfb = c_loc(wfb)
        d%amph = c_loc(amph); d%acph = c_loc(acph); d%apph = c_loc(apph)
        d%amvr = c_loc(amvr); d%acvr = c_loc(acvr); d%apvr = c_loc(apvr)
        d%ak1 = c_loc(ak1); d%ak3 = c_loc(ak3)
        call chs_check(c_create(d, c_null_ptr, handle), 'chs_init')
    end subroutine

    subroutine chs_finalize()
        call chs_check(c_destroy(handle), 'chs_finalize')
        handle = c_null_ptr
    end subroutine

    !> Q_tg(NCL1,0:NCL2+1,NCL3,3), PR_tg(NCL1,0:NCL2+1,NCL3), CONVH0_tg(NCL1,NCL2,NCL3,3)
    subroutine chs_put_state(q, pr, convh0)
        real(c_double), target, contiguous, intent(in) :: q(:, :, :, :), pr(:, :, :), convh0(:, :, :, :)
        call chs_check(c_upload(handle, c_loc(q), c_loc(pr), c_loc(convh0)), 'chs_put_state')
    end subroutine

    subroutine chs_get_state(q, pr, convh0)
        real(c_double), target, contiguous, intent(inout) :: q(:, :, :, :), pr(:, :, :), convh0(:, :, :, :)
        call chs_check(c_download(handle, c_loc(q), c_loc(pr), c_loc(convh0)), 'chs_get_state')
    end subroutine

    !> SOLVERRK3_MOM_tg(NS) on the device-resident state (src/SOLVERRK3_MOM_tg.f90:18)
    subroutine chs_solverrk3_mom_tg(ns, dt)
        integer, intent(in) :: ns
        real(c_double), intent(in) :: dt
        call chs_check(c_substep(handle, int(ns, c_int), dt), 'chs_solverrk3_mom_tg')
    end subroutine

    !> DIVGCK_tg: max|div u| (src/DIVGCK.f90:16)
    subroutine chs_divgck_tg(maxdiv)
        real(c_double), intent(out) :: maxdiv
        call chs_check(c_divgck(handle, maxdiv), 'chs_divgck_tg')
    end subroutine

    !> CFL_tg: max convective rate (src/CFL.f90:21)
    subroutine chs_cfl_tg(rate)
        real(c_double), intent(out) :: rate
        call chs_check(c_cfl(handle, rate), 'chs_cfl_tg')
    end subroutine

    subroutine chs_tdmaijj_noncyc(a, b, c, r, bcj, js, jsz, is, isz)
        integer, intent(in) :: js, jsz, is, isz
        real(c_double), target, intent(in) :: a(is:isz + is - 1, js:jsz + js - 1), b(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(in) :: c(is:isz + is - 1, js:jsz + js - 1), bcj(is:isz + is - 1, 2)
        real(c_double), target, intent(inout) :: r(is:isz + is - 1, js:jsz + js - 1)
        call chs_check(c_tdma_host(c_loc(a), c_loc(b), c_loc(c), c_loc(r), c_loc(bcj), int(jsz, c_int), &
                       int(isz, c_int), int(isz, c_long_long), 1_c_long_long, 0_c_int), 'TDMAIJJ_nonCYC')
    end subroutine

    subroutine chs_tdmaiji_noncyc(a, b, c, r, bci, is, isz, js, jsz)
        integer, intent(in) :: is, isz, js, jsz
        real(c_double), target, intent(in) :: a(is:isz + is - 1, js:jsz + js - 1), b(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(in) :: c(is:isz + is - 1, js:jsz + js - 1), bci(js:jsz + js - 1, 2)
        real(c_double), target, intent(inout) :: r(is:isz + is - 1, js:jsz + js - 1)
        call chs_check(c_tdma_host(c_loc(a), c_loc(b), c_loc(c), c_loc(r), c_loc(bci), int(isz, c_int), &
                       int(jsz, c_int), 1_c_long_long, int(isz, c_long_long), 0_c_int), 'TDMAIJI_nonCYC')
    end subroutine

    subroutine chs_tdmaijj_cyc(a, b, c, r, js, jsz, is, isz)
        integer, intent(in) :: js, jsz, is, isz
        real(c_double), target, intent(in) :: a(is:isz + is - 1, js:jsz + js - 1), b(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(in) :: c(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(inout) :: r(is:isz + is - 1, js:jsz + js - 1)
        call chs_check(c_tdma_host(c_loc(a), c_loc(b), c_loc(c), c_loc(r), c_null_ptr, int(jsz, c_int), &
                       int(isz, c_int), int(isz, c_long_long), 1_c_long_long, 1_c_int), 'TDMAIJJ_CYC')
    end subroutine

    subroutine chs_tdmaiji_cyc(a, b, c, r, is, isz, js, jsz)
        integer, intent(in) :: is, isz, js, jsz
        real(c_double), target, intent(in) :: a(is:isz + is - 1, js:jsz + js - 1), b(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(in) :: c(is:isz + is - 1, js:jsz + js - 1)
        real(c_double), target, intent(inout) :: r(is:isz + is - 1, js:jsz + js - 1)
        call chs_check(c_tdma_host(c_loc(a), c_loc(b), c_loc(c), c_loc(r), c_null_ptr, int(isz, c_int), &
                       int(jsz, c_int), 1_c_long_long, int(isz, c_long_long), 1_c_int), 'TDMAIJI_CYC')
    end subroutine

end module chapsim_gpu
