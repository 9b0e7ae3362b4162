!> Drop-in replacements under the reference's own names and argument lists.
!>
!> Compile this file into a CHAPSim build in place of the reference sources
!> it replaces (src/TDMAIJ{I,J}_{CYC,nonCYC}.f90, the FFT99_POIS3D_periodicxz
!> part of src/possion3d_FFT99.f90, SOLVERRK3_MOM_tg in
!> src/SOLVERRK3_MOM_tg.f90, THERM_PROP_UPDATE in src/THERM_PROP_UPDATE.f90,
!> src/SOLVERRK3_ENG_IO.f90 and src/SOLVERRK3_MOM_IO.f90), with the same
!> compiler and after fortran/chapsim_gpu.f90.  The routines read and write
!> the reference's module arrays (modules cparam, init_info, flow_info,
!> thermal_info of src/modules.f90) exactly where the originals do; the call
!> sites stay unchanged.
!>
!> Before the time loop the driver calls, once:
!>     CALL CHS_INIT(...)             (mesh and coefficients, fortran/chapsim_gpu.f90)
!>     CALL CHS_PROPS_INIT(...)       heated channel: the scaled table
!>     CALL CHS_HEATED_INIT(...)      heated channel: device mirrors and io switches
!>
!> Each call here moves its module arrays to the GPU and back (a routine-level
!> drop-in: PCIe per call).  The device-resident calls of module chapsim_gpu
!> (CHS_PUT_STATE / CHS_SOLVERRK3_MOM_TG / CHS_GET_STATE and CHS_HEATED_PUT /
!> CHS_SOLVERRK3_ENG_IO / ... / CHS_HEATED_GET) keep the fields in HBM
!> between substeps instead.

!> TDMAIJJ_nonCYC(A,B,C,R,BCJ,JS,JSZ,IS,ISZ) (src/TDMAIJJ_nonCYC.f90:67)
subroutine TDMAIJJ_nonCYC(A, B, C, R, BCJ, JS, JSZ, IS, ISZ)
    use chapsim_gpu, only: chs_tdmaijj_noncyc
    implicit none
    integer(4), intent(in) :: JS, JSZ, IS, ISZ
    real(8), intent(in) :: A(IS:ISZ + IS - 1, JS:JSZ + JS - 1), B(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(in) :: C(IS:ISZ + IS - 1, JS:JSZ + JS - 1), BCJ(IS:ISZ + IS - 1, 2)
    real(8), intent(inout) :: R(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    call chs_tdmaijj_noncyc(A, B, C, R, BCJ, JS, JSZ, IS, ISZ)
end subroutine

!> TDMAIJI_nonCYC(A,B,C,R,BCI,IS,ISZ,JS,JSZ) (src/TDMAIJI_nonCYC.f90:67)
subroutine TDMAIJI_nonCYC(A, B, C, R, BCI, IS, ISZ, JS, JSZ)
    use chapsim_gpu, only: chs_tdmaiji_noncyc
    implicit none
    integer(4), intent(in) :: IS, ISZ, JS, JSZ
    real(8), intent(in) :: A(IS:ISZ + IS - 1, JS:JSZ + JS - 1), B(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(in) :: C(IS:ISZ + IS - 1, JS:JSZ + JS - 1), BCI(JS:JSZ + JS - 1, 2)
    real(8), intent(inout) :: R(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    call chs_tdmaiji_noncyc(A, B, C, R, BCI, IS, ISZ, JS, JSZ)
end subroutine

!> TDMAIJJ_CYC(A,B,C,R,JS,JSZ,IS,ISZ) (src/TDMAIJJ_CYC.f90:67)
subroutine TDMAIJJ_CYC(A, B, C, R, JS, JSZ, IS, ISZ)
    use chapsim_gpu, only: chs_tdmaijj_cyc
    implicit none
    integer(4), intent(in) :: JS, JSZ, IS, ISZ
    real(8), intent(in) :: A(IS:ISZ + IS - 1, JS:JSZ + JS - 1), B(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(in) :: C(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(inout) :: R(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    call chs_tdmaijj_cyc(A, B, C, R, JS, JSZ, IS, ISZ)
end subroutine

!> TDMAIJI_CYC(A,B,C,R,IS,ISZ,JS,JSZ) (src/TDMAIJI_CYC.f90:67)
subroutine TDMAIJI_CYC(A, B, C, R, IS, ISZ, JS, JSZ)
    use chapsim_gpu, only: chs_tdmaiji_cyc
    implicit none
    integer(4), intent(in) :: IS, ISZ, JS, JSZ
    real(8), intent(in) :: A(IS:ISZ + IS - 1, JS:JSZ + JS - 1), B(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(in) :: C(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    real(8), intent(inout) :: R(IS:ISZ + IS - 1, JS:JSZ + JS - 1)
    call chs_tdmaiji_cyc(A, B, C, R, IS, ISZ, JS, JSZ)
end subroutine

!> FFT99_POIS3D_periodicxz(IDOMAIN) (src/possion3d_FFT99.f90:254-488):
!> RHSLLPHI_tg -> DPH_tg (IDOMAIN = ITG) or RHSLLPHI_io -> DPH_io (IIO).
subroutine FFT99_POIS3D_periodicxz(IDOMAIN)
    use cparam, only: ITG
    use flow_info, only: RHSLLPHI_tg, DPH_tg, RHSLLPHI_io, DPH_io
    use chapsim_gpu, only: chs_fft99_pois3d
    implicit none
    integer(4), intent(in) :: IDOMAIN
    if (IDOMAIN == ITG) then
        call chs_fft99_pois3d(RHSLLPHI_tg, DPH_tg)
    else
        call chs_fft99_pois3d(RHSLLPHI_io, DPH_io)
    end if
end subroutine

!> SOLVERRK3_MOM_tg(NS) (src/SOLVERRK3_MOM_tg.f90:18-61) on Q_tg, PR_tg,
!> CONVH0_tg with the time step DT of module init_info.
subroutine SOLVERRK3_MOM_tg(NS)
    use init_info, only: DT
    use flow_info, only: Q_tg, PR_tg, CONVH0_tg
    use chapsim_gpu, only: chs_put_state, chs_solverrk3_mom_tg, chs_get_state
    implicit none
    integer(4), intent(in) :: NS
    call chs_put_state(Q_tg, PR_tg, CONVH0_tg)
    call chs_solverrk3_mom_tg(NS, DT)
    call chs_get_state(Q_tg, PR_tg, CONVH0_tg)
end subroutine

!> THERM_PROP_UPDATE(IREGION) (src/THERM_PROP_UPDATE.f90:129-205) for the
!> periodic main domain (IREGION = IMAIND, I = 1..NCL1_io): TEMPERATURE,
!> DENSITY, VISCOUSITY, THERMCONDT of rows 1..N2DO from ENTHALPY.
subroutine THERM_PROP_UPDATE(IREGION)
    use thermal_info, only: ENTHALPY, TEMPERATURE, DENSITY, VISCOUSITY, THERMCONDT
    use chapsim_gpu, only: chs_therm_prop_update
    implicit none
    integer(4), intent(in) :: IREGION
    call chs_therm_prop_update(ENTHALPY, TEMPERATURE, DENSITY, VISCOUSITY, THERMCONDT)
end subroutine

!> The heated channel's module arrays to the device mirrors and back.
subroutine chs_dropin_heated(dir)
    use flow_info, only: Q_io, G_io, G0_io, PR_io, EXPLT0_io, DivU_io
    use thermal_info, only: RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, &
                            THERMCONDT, RHS_ENERGY0, WALLFLUX, DrhoDtP, RHO_STG, MU_STG, DRHOI_STG
    use chapsim_gpu, only: chs_heated_put, chs_heated_get
    implicit none
    integer, intent(in) :: dir
    if (dir == 1) then
        call chs_heated_put(RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, &
                            THERMCONDT, RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, &
                            MU_STG, EXPLT0_io, DRHOI_STG)
    else
        call chs_heated_get(RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0, DENSITY1, VISCOUSITY, VISCOUSITY0, &
                            THERMCONDT, RHS_ENERGY0, G_io, G0_io, WALLFLUX, Q_io, PR_io, DrhoDtP, DivU_io, RHO_STG, &
                            MU_STG, EXPLT0_io, DRHOI_STG)
    end if
end subroutine

!> SOLVERRK3_ENG_IO(NS) (src/SOLVERRK3_ENG_IO.f90:1-77) with DT (init_info)
!> and CTHECD (thermal_info).
subroutine SOLVERRK3_ENG_IO(NS)
    use init_info, only: DT
    use thermal_info, only: CTHECD
    use chapsim_gpu, only: chs_solverrk3_eng_io
    implicit none
    integer(4), intent(in) :: NS
    call chs_dropin_heated(1)
    call chs_solverrk3_eng_io(NS, DT, CTHECD)
    call chs_dropin_heated(2)
end subroutine

!> SOLVERRK3_MOM_IO(NS) (src/SOLVERRK3_MOM_IO.f90:1-168) with DT (init_info),
!> after the driver's own DENSITY_TIME_DERIVATION, DENSITY_Staggered,
!> MU_Staggered and VELOCITY_CALC_io (src/SOLVE.f90:121-133) have set
!> DrhoDtP, RHO_STG, MU_STG and Q_io; the io switches (NFLOW, FLOWDRVTP,
!> GRAVDIR, visthemflg, ...) are those given to CHS_HEATED_INIT.
subroutine SOLVERRK3_MOM_IO(NS)
    use init_info, only: DT
    use chapsim_gpu, only: chs_solverrk3_mom_io
    implicit none
    integer(4), intent(in) :: NS
    call chs_dropin_heated(1)
    call chs_solverrk3_mom_io(NS, DT)
    call chs_dropin_heated(2)
end subroutine
