!> The module variables the drop-ins of chapsim_dropin.f90 read and write,
!> declared under the reference's names and types (src/modules.f90: cparam
!> :22-100, init_info :229-307, flow_info :310-374, thermal_info :377-520).
!> For the test program test_dropin.f90 only: a CHAPSim build compiles the
!> drop-ins against the reference's own src/modules.f90 instead.
module cparam
    implicit none
    integer, parameter :: WP = kind(1.0d0)
    integer(4), parameter :: ITG = 1
    integer(4), parameter :: IIO = 2
end module cparam

module init_info
    use cparam
    implicit none
    real(WP) :: DT
end module init_info

module flow_info
    use cparam
    implicit none
    real(WP), allocatable :: Q_tg(:, :, :, :), PR_tg(:, :, :), CONVH0_tg(:, :, :, :)
    real(WP), allocatable :: DPH_tg(:, :, :), RHSLLPHI_tg(:, :, :)
    real(WP), allocatable :: Q_io(:, :, :, :), G_io(:, :, :, :), G0_io(:, :, :, :), PR_io(:, :, :)
    real(WP), allocatable :: EXPLT0_io(:, :, :, :), DPH_io(:, :, :), RHSLLPHI_io(:, :, :), DivU_io(:, :, :)
end module flow_info

module thermal_info
    use cparam
    implicit none
    real(WP) :: CTHECD
    real(WP), allocatable :: RHOH(:, :, :), DENSITY(:, :, :), TEMPERATURE(:, :, :), VISCOUSITY(:, :, :)
    real(WP), allocatable :: VISCOUSITY0(:, :, :), THERMCONDT(:, :, :), ENTHALPY(:, :, :)
    real(WP), allocatable :: RHO_STG(:, :, :, :), DRHOI_STG(:, :, :, :), MU_STG(:, :, :, :)
    real(WP), allocatable :: RHS_ENERGY0(:, :, :), DENSITY0(:, :, :), DrhoDtP(:, :, :), WALLFLUX(:, :)
    real(WP), allocatable :: DENSITY1(:, :, :)
end module thermal_info
