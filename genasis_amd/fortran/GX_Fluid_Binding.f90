!-- GX_Fluid_Binding: ISO_C_BINDING interface to libgx_fluid.so (include/gx_fluid.h)
!   for a GenASiS host.  Step_RK_C_ASC_Template % Compute and
!   Integrator_C_PS % ComputeTimeStep_C_ASC call gx_step_rk2 and
!   gx_compute_time_step when StepBackend = 'HIP'.

module GX_Fluid_Binding

  use iso_c_binding
  implicit none
  private

  integer ( c_int ), public, parameter :: &
    GX_OK = 0, &
    GX_BC_PERIODIC = 0, GX_BC_OUTFLOW = 1, GX_BC_REFLECTING = 2, &
    GX_HLL = 0, GX_HLLC = 1, &
    GX_FP64 = 0, GX_FP32 = 1, &
    GX_N_FIELDS = 23

  type, public, bind ( C ) :: gx_grid
    integer ( c_int ) :: nCells ( 3 )
    real ( c_double ) :: MinCoordinate ( 3 )
    real ( c_double ) :: MaxCoordinate ( 3 )
    integer ( c_int ) :: BoundaryCondition ( 6 )
  end type gx_grid

  type, public, bind ( C ) :: gx_params
    real ( c_double ) :: AdiabaticIndex
    real ( c_double ) :: LimiterParameter
    real ( c_double ) :: ShockThreshold
    real ( c_double ) :: CourantFactor
    real ( c_double ) :: BaryonDensityMin
    real ( c_double ) :: BaryonMassReference
    integer ( c_int ) :: UseLimiter
    integer ( c_int ) :: RiemannSolverType
    integer ( c_int ) :: Precision
    integer ( c_int ) :: ReconstructionType
  end type gx_params

  public :: &
    gx_default_params, gx_create, gx_destroy, gx_upload, gx_download, &
    gx_compute_time_step, gx_step_rk2, gx_sync, gx_totals, gx_last_error, &
    gx_compute_tally, gx_get_boundary_fluence, gx_write_checkpoint, &
    gx_read_checkpoint, gx_set_record_sources, gx_get_sources, &
    LastErrorString

  !-- gx_read_checkpoint modes: the reference's RestartFrom, or exact
  integer ( c_int ), parameter, public :: &
    GX_RESTART_PRIMITIVE = 0, GX_RESTART_EXACT = 1

  !-- Tally_F_P % Value (12): BaryonNumber, Momentum_1..3, KineticEnergy,
  !   AngularMomentum_1..3, GravitationalEnergy, TotalEnergy, FluidEnergy,
  !   InternalEnergy
  integer ( c_int ), parameter, public :: GX_TALLY_N = 12

  interface

    subroutine gx_default_params ( p ) bind ( C, name = 'gx_default_params' )
      import :: gx_params
      type ( gx_params ), intent ( out ) :: p
    end subroutine gx_default_params

    integer ( c_int ) function gx_create ( grid, params, ctx ) &
                        bind ( C, name = 'gx_create' )
      import :: c_int, c_ptr, gx_grid, gx_params
      type ( gx_grid ), intent ( in ) :: grid
      type ( gx_params ), intent ( in ) :: params
      type ( c_ptr ), intent ( out ) :: ctx
    end function gx_create

    integer ( c_int ) function gx_destroy ( ctx ) bind ( C, name = 'gx_destroy' )
      import :: c_int, c_ptr
      type ( c_ptr ), value :: ctx
    end function gx_destroy

    !-- Value ( nValues, nVariables ) of the Current, column-major;
    !   FieldMap ( GX_N_FIELDS ) = column of each Fluid_P_I field (1-based)
    integer ( c_int ) function gx_upload ( ctx, Value, ldValue, FieldMap ) &
                        bind ( C, name = 'gx_upload' )
      import :: c_int, c_ptr, c_double, c_long_long
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( in ) :: Value ( * )
      integer ( c_long_long ), value :: ldValue
      integer ( c_int ), intent ( in ) :: FieldMap ( * )
    end function gx_upload

    integer ( c_int ) function gx_download ( ctx, Value, ldValue, FieldMap ) &
                        bind ( C, name = 'gx_download' )
      import :: c_int, c_ptr, c_double, c_long_long
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( inout ) :: Value ( * )
      integer ( c_long_long ), value :: ldValue
      integer ( c_int ), intent ( in ) :: FieldMap ( * )
    end function gx_download

    integer ( c_int ) function gx_compute_time_step ( ctx, Candidate ) &
                        bind ( C, name = 'gx_compute_time_step' )
      import :: c_int, c_ptr, c_double
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( out ) :: Candidate
    end function gx_compute_time_step

    integer ( c_int ) function gx_step_rk2 ( ctx, Time, TimeStep ) &
                        bind ( C, name = 'gx_step_rk2' )
      import :: c_int, c_ptr, c_double
      type ( c_ptr ), value :: ctx
      real ( c_double ), value :: Time, TimeStep
    end function gx_step_rk2

    integer ( c_int ) function gx_sync ( ctx ) bind ( C, name = 'gx_sync' )
      import :: c_int, c_ptr
      type ( c_ptr ), value :: ctx
    end function gx_sync

    integer ( c_int ) function gx_totals ( ctx, Total ) &
                        bind ( C, name = 'gx_totals' )
      import :: c_int, c_ptr, c_double
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( out ) :: Total ( 5 )
    end function gx_totals

    !-- Current_ASC % ComputeTally for the chart: Interior, Boundary, Total
    !   and Change, GX_TALLY_N values each
    integer ( c_int ) function gx_compute_tally &
                        ( ctx, ComputeChange, Interior, Boundary, Total, Change ) &
                        bind ( C, name = 'gx_compute_tally' )
      import :: c_int, c_ptr, c_double, GX_TALLY_N
      type ( c_ptr ), value :: ctx
      integer ( c_int ), value :: ComputeChange
      real ( c_double ), intent ( out ) :: &
        Interior ( GX_TALLY_N ), Boundary ( GX_TALLY_N ), &
        Total ( GX_TALLY_N ), Change ( GX_TALLY_N )
    end function gx_compute_tally

    !-- BoundaryFluence_CSL of the last step: faces 1..6 (inner and outer of
    !   each dimension) x D, S_1, S_2, S_3, G x the face's slab of proper cells
    integer ( c_int ) function gx_get_boundary_fluence ( ctx, Fluence ) &
                        bind ( C, name = 'gx_get_boundary_fluence' )
      import :: c_int, c_ptr, c_double
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( out ) :: Fluence ( * )
    end function gx_get_boundary_fluence

    !-- Current % Sources (RecordDivergence), off until switched on; D,
    !   S_1..3, G of the proper cells, x fastest
    integer ( c_int ) function gx_set_record_sources ( ctx, On ) &
                        bind ( C, name = 'gx_set_record_sources' )
      import :: c_int, c_ptr
      type ( c_ptr ), value :: ctx
      integer ( c_int ), value :: On
    end function gx_set_record_sources

    integer ( c_int ) function gx_get_sources ( ctx, Sources ) &
                        bind ( C, name = 'gx_get_sources' )
      import :: c_int, c_ptr, c_double
      type ( c_ptr ), value :: ctx
      real ( c_double ), intent ( out ) :: Sources ( * )
    end function gx_get_sources

    !-- Checkpoint file at AdministerCheckpoint; Path is a NUL-terminated
    !   character array (trim ( Path ) // c_null_char)
    integer ( c_int ) function gx_write_checkpoint &
                        ( ctx, Path, Time, iCycle, iCheckpoint ) &
                        bind ( C, name = 'gx_write_checkpoint' )
      import :: c_int, c_ptr, c_double, c_char, c_long_long
      type ( c_ptr ), value :: ctx
      character ( kind = c_char ), intent ( in ) :: Path ( * )
      real ( c_double ), value :: Time
      integer ( c_long_long ), value :: iCycle
      integer ( c_int ), value :: iCheckpoint
    end function gx_write_checkpoint

    !-- RestartFrom: Time, cycle and iCheckpoint of the file
    integer ( c_int ) function gx_read_checkpoint &
                        ( ctx, Path, Mode, Time, iCycle, iCheckpoint ) &
                        bind ( C, name = 'gx_read_checkpoint' )
      import :: c_int, c_ptr, c_double, c_char, c_long_long
      type ( c_ptr ), value :: ctx
      character ( kind = c_char ), intent ( in ) :: Path ( * )
      integer ( c_int ), value :: Mode
      real ( c_double ), intent ( out ) :: Time
      integer ( c_long_long ), intent ( out ) :: iCycle
      integer ( c_int ), intent ( out ) :: iCheckpoint
    end function gx_read_checkpoint

    type ( c_ptr ) function gx_last_error ( ) bind ( C, name = 'gx_last_error' )
      import :: c_ptr
    end function gx_last_error

  end interface

contains

  function LastErrorString ( ) result ( S )
    character ( len = : ), allocatable :: S
    character ( kind = c_char ), pointer :: C ( : )
    integer :: n
    type ( c_ptr ) :: p
    p = gx_last_error ( )
    call c_f_pointer ( p, C, [ 4096 ] )
    n = 0
    do while ( n < 4096 )
      if ( C ( n + 1 ) == c_null_char ) exit
      n = n + 1
    end do
    allocate ( character ( len = n ) :: S )
    S = transfer ( C ( 1 : n ), S )
  end function LastErrorString

end module GX_Fluid_Binding
