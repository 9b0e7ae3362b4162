!> Calls the library through the Fortran shim exactly as a CHAPSim host
!> program would: the four TDMA drop-ins on host arrays with the reference's
!> argument lists, checked against a plain Thomas / residual computed here.
program test_shim
    use iso_c_binding
    use chapsim_gpu
    implicit none
    integer, parameter :: isz = 37, jsz = 24
    real(c_double) :: a(isz, jsz), b(isz, jsz), c(isz, jsz), r(isz, jsz), d(isz, jsz)
    real(c_double) :: bcj(isz, 2), bci(jsz, 2), res, worst
    integer :: i, j, k
    worst = 0.0d0
    do j = 1, jsz
        do i = 1, isz
            a(i, j) = -0.1d0 - 0.8d0 * modulo(0.37d0 * i + 0.61d0 * j, 1.0d0)
            c(i, j) = -0.1d0 - 0.8d0 * modulo(0.23d0 * i + 0.47d0 * j, 1.0d0)
            b(i, j) = 2.0d0 + modulo(0.11d0 * i * j, 1.0d0)
            d(i, j) = sin(0.3d0 * i) + cos(0.7d0 * j)
        end do
    end do
    do i = 1, isz
        bcj(i, 1) = 0.5d0 * i / isz; bcj(i, 2) = -0.25d0
    end do
    do j = 1, jsz
        bci(j, 1) = 0.1d0; bci(j, 2) = 0.2d0 * j / jsz
    end do
    ! 1. TDMAIJJ_nonCYC: lines i, march along j
    r = d
    call chs_tdmaijj_noncyc(a, b, c, r, bcj, 1, jsz, 1, isz)
    do i = 1, isz
        do j = 1, jsz
            res = b(i, j) * r(i, j) - d(i, j)
            if (j > 1) then
                res = res + a(i, j) * r(i, j - 1)
            else
                res = res + a(i, j) * bcj(i, 1)
            end if
            if (j < jsz) then
                res = res + c(i, j) * r(i, j + 1)
            else
                res = res + c(i, j) * bcj(i, 2)
            end if
            worst = max(worst, abs(res))
        end do
    end do
    ! 2. TDMAIJI_nonCYC: lines j, march along i
    r = d
    call chs_tdmaiji_noncyc(a, b, c, r, bci, 1, isz, 1, jsz)
    do j = 1, jsz
        do i = 1, isz
            res = b(i, j) * r(i, j) - d(i, j)
            if (i > 1) then
                res = res + a(i, j) * r(i - 1, j)
            else
                res = res + a(i, j) * bci(j, 1)
            end if
            if (i < isz) then
                res = res + c(i, j) * r(i + 1, j)
            else
                res = res + c(i, j) * bci(j, 2)
            end if
            worst = max(worst, abs(res))
        end do
    end do
    ! 3./4. cyclic forms
    r = d
    call chs_tdmaijj_cyc(a, b, c, r, 1, jsz, 1, isz)
    do i = 1, isz
        do j = 1, jsz
            k = j - 1; if (k < 1) k = jsz
            res = a(i, j) * r(i, k) + b(i, j) * r(i, j) - d(i, j)
            k = j + 1; if (k > jsz) k = 1
            res = res + c(i, j) * r(i, k)
            worst = max(worst, abs(res))
        end do
    end do
    r = d
    call chs_tdmaiji_cyc(a, b, c, r, 1, isz, 1, jsz)
    do j = 1, jsz
        do i = 1, isz
            k = i - 1; if (k < 1) k = isz
            res = a(i, j) * r(k, j) + b(i, j) * r(i, j) - d(i, j)
            k = i + 1; if (k > isz) k = 1
            res = res + c(i, j) * r(k, j)
            worst = max(worst, abs(res))
        end do
    end do
    write(*, '(A,ES12.4)') 'fortran shim: max residual of the four TDMA drop-ins = ', worst
    if (worst > 1.0d-13) error stop 2
end program test_shim
