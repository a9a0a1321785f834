!-- Host-only check of the binding: default parameters and the argument
!   errors that gx_create reports before touching a device.
program GX_Binding_Check
  use iso_c_binding
  use GX_Fluid_Binding
  implicit none
  type ( gx_grid ) :: G
  type ( gx_params ) :: P
  type ( c_ptr ) :: ctx
  integer ( c_int ) :: rc

  call gx_default_params ( P )
  if ( P % AdiabaticIndex /= 1.4_c_double .or. P % CourantFactor /= 0.7_c_double &
       .or. P % RiemannSolverType /= GX_HLLC .or. P % UseLimiter /= 1 ) stop 1

  G % nCells = [ 1, 8, 8 ]
  G % MinCoordinate = 0.0_c_double
  G % MaxCoordinate = 1.0_c_double
  G % BoundaryCondition = GX_BC_OUTFLOW
  rc = gx_create ( G, P, ctx )
  if ( rc /= -1 ) stop 2
  if ( index ( LastErrorString ( ), 'nCells' ) == 0 ) stop 3
  print '(a)', 'GX_Binding_Check: ' // LastErrorString ( )
end program GX_Binding_Check
