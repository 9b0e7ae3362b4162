!> Drives the reference-named drop-ins (chapsim_dropin.f90) as a CHAPSim
!> program would, on module arrays (test_modules.f90): reads the case from a
!> stream file written by tests/test_gpu_fortran.py, runs
!>   TDMAIJJ_nonCYC, TDMAIJI_nonCYC, TDMAIJJ_CYC, TDMAIJI_CYC on small systems,
!>   FFT99_POIS3D_periodicxz(ITG),
!>   SOLVERRK3_MOM_tg(NS) for NS = 0, 1, 2, 3,
!>   THERM_PROP_UPDATE, SOLVERRK3_ENG_IO(1), the io staging on the device
!>   (CHS_HEATED_PUT / CHS_IO_PREPARE / CHS_HEATED_GET) and SOLVERRK3_MOM_IO(1),
!> and writes every result to a stream file for the test to compare with the
!> oracle.
!>     test_dropin <input> <output>
program test_dropin
    use iso_c_binding
    use cparam
    use init_info
    use flow_info
    use thermal_info
    use chapsim_gpu
    implicit none
    integer(4) :: ni(9), L, M, N, ntab, ns, tl, tm
    real(8) :: nr(11), dt_tg, dt_io
    real(8), allocatable :: dyfi(:), dyci(:), wff(:), wfb(:), amph(:), acph(:), apph(:)
    real(8), allocatable :: amvr(:, :), acvr(:, :), apvr(:, :), ak1(:), ak3(:)
    real(8), allocatable :: th(:), tt(:), td(:), tmu(:), tk(:), tcp(:)
    real(8), allocatable :: ta(:, :), tb(:, :), tc(:, :), tr(:, :), tbcj(:, :), tbci(:, :)
    character(len=512) :: fin, fout
    integer :: u, v, q
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    open(newunit=v, file=trim(fout), access='stream', form='unformatted', status='replace')
    read(u) ni, nr
    L = ni(1); M = ni(2); N = ni(3); ntab = ni(4)
    dt_tg = nr(7); dt_io = nr(8)
    allocate(dyfi(M), dyci(M + 1), wff(M + 1), wfb(M + 1), amph(M), acph(M), apph(M))
    allocate(amvr(M + 1, 3), acvr(M + 1, 3), apvr(M + 1, 3), ak1(L), ak3(N))
    read(u) dyfi, dyci, wff, wfb, amph, acph, apph, amvr, acvr, apvr, ak1, ak3
    call chs_init(L, M, N, nr(1), nr(2), nr(3), nr(4), nr(5), nr(6), dyfi, dyci, wff, wfb, amph, acph, apph, &
                  amvr, acvr, apvr, ak1, ak3, CHS_FP64)

    ! the four line solvers on (TL, TM) systems with lower bounds (2, 0)
    read(u) tl, tm
    allocate(ta(2:tl + 1, 0:tm - 1), tb(2:tl + 1, 0:tm - 1), tc(2:tl + 1, 0:tm - 1), tr(2:tl + 1, 0:tm - 1))
    allocate(tbcj(2:tl + 1, 2), tbci(0:tm - 1, 2))
    do q = 1, 4
        read(u) ta, tb, tc, tr
        select case (q)
        case (1)
            read(u) tbcj
            call TDMAIJJ_nonCYC(ta, tb, tc, tr, tbcj, 0, tm, 2, tl)
        case (2)
            read(u) tbci
            call TDMAIJI_nonCYC(ta, tb, tc, tr, tbci, 2, tl, 0, tm)
        case (3)
            call TDMAIJJ_CYC(ta, tb, tc, tr, 0, tm, 2, tl)
        case (4)
            call TDMAIJI_CYC(ta, tb, tc, tr, 2, tl, 0, tm)
        end select
        write(v) tr
    end do

    ! Poisson
    allocate(RHSLLPHI_tg(L, M, N), DPH_tg(L, 0:M + 1, N))
    read(u) RHSLLPHI_tg
    DPH_tg = 7.0d0
    call FFT99_POIS3D_periodicxz(ITG)
    write(v) DPH_tg

    ! tg substeps
    allocate(Q_tg(L, 0:M + 1, N, 3), PR_tg(L, 0:M + 1, N), CONVH0_tg(L, M, N, 3))
    read(u) Q_tg, PR_tg, CONVH0_tg
    DT = dt_tg
    do ns = 0, 3
        call SOLVERRK3_MOM_tg(ns)
    end do
    write(v) Q_tg, PR_tg, CONVH0_tg

    ! heated channel
    allocate(th(ntab), tt(ntab), td(ntab), tmu(ntab), tk(ntab), tcp(ntab))
    read(u) th, tt, td, tmu, tk, tcp
    call chs_props_init(ntab, th, tt, td, tmu, tk, tcp)
    allocate(RHOH(L, 0:M + 1, N), ENTHALPY(L, 0:M + 1, N), TEMPERATURE(L, 0:M + 1, N), DENSITY(L, 0:M + 1, N))
    allocate(DENSITY0(L, 0:M + 1, N), DENSITY1(L, 0:M + 1, N), VISCOUSITY(L, 0:M + 1, N))
    allocate(VISCOUSITY0(L, 0:M + 1, N), THERMCONDT(L, 0:M + 1, N), RHS_ENERGY0(L, M, N))
    allocate(G_io(L, 0:M + 1, N, 3), G0_io(L, 0:M + 1, N, 3), WALLFLUX(L, 2), Q_io(L, 0:M + 1, N, 3))
    allocate(PR_io(L, 0:M + 1, N), DrhoDtP(L, 0:M + 1, N), DivU_io(L, 0:M + 1, N), RHO_STG(L, 0:M + 1, N, 3))
    allocate(MU_STG(L, 0:M + 1, N, 3), EXPLT0_io(L, M, N, 3), DRHOI_STG(L, 0:M + 1, N, 3))
    read(u) RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, THERMCONDT, &
            RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, MU_STG, EXPLT0_io, DRHOI_STG
    close(u)
    ! THERM_PROP_UPDATE on the start state, written, then the start state back
    block
        real(8), allocatable :: t0(:, :, :), d0(:, :, :), m0(:, :, :), k0(:, :, :)
        t0 = TEMPERATURE; d0 = DENSITY; m0 = VISCOUSITY; k0 = THERMCONDT
        TEMPERATURE = -1.0d0; DENSITY = -1.0d0; VISCOUSITY = -1.0d0; THERMCONDT = -1.0d0
        call THERM_PROP_UPDATE(3)
        write(v) TEMPERATURE, DENSITY, VISCOUSITY, THERMCONDT
        TEMPERATURE = t0; DENSITY = d0; VISCOUSITY = m0; THERMCONDT = k0
    end block
    CTHECD = nr(9)
    DT = dt_io
    call chs_heated_init(ni(5), ni(6), ni(7), ni(8), nr(10), nr(11), ni(9))
    call SOLVERRK3_ENG_IO(1)
    call chs_heated_put(RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, &
                        THERMCONDT, RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, &
                        MU_STG, EXPLT0_io, DRHOI_STG)
    call chs_io_prepare(1)
    call chs_heated_get(RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, &
                        THERMCONDT, RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, &
                        MU_STG, EXPLT0_io, DRHOI_STG)
    call SOLVERRK3_MOM_IO(1)
    write(v) RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, THERMCONDT, &
             RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, MU_STG, EXPLT0_io, DRHOI_STG
    close(v)
    call chs_finalize()
    write(*, '(A)') 'test_dropin: done'
end program test_dropin
