! mrmg_m.f90 -- Fortran 2003 binding of the MI355X time-stepping library (include/mrmg.h).
!
! This is the host side a Fortran driver uses to hand its `data_m`/`field_m` arrays to the
! GPU path.  The derived types are interoperable copies of the C structs; every entry
! point returns an integer status (0 = ok), the Fortran form of the reference's
! `logical ok_p` idiom (cli.f90:249), and mrmg_error() gives the message.
!
! The handle convention follows the reference's own C-ABI precedent for its C++ module
! (src/FP/fp.f90:12-44): an opaque type(c_ptr) passed by value, scalars by reference.
module mrmg_m
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: mrmg_geom, mrmg_layer, mrmg_ufield, mrmg_source, mrmg_step_params, mrmg_diag, mrmg_sensor
  public :: mrmg_create, mrmg_create_open, mrmg_destroy, mrmg_set_tensor, mrmg_build_tensor, mrmg_build_tensor_images, &
            mrmg_set_layer, &
            mrmg_set_ufield, mrmg_set_state, mrmg_get_state, mrmg_set_mag, mrmg_get_mag, mrmg_step, mrmg_error, &
            mrmg_readback, mrmg_switch_arm, mrmg_switch_read, mrmg_set_source, mrmg_tune, mrmg_get_tune, &
            mrmg_philox_normals, mrmg_footprint
  public :: MRMG_FP64, MRMG_FP32, MRMG_NOISE_NONE, MRMG_NOISE_INJECT, MRMG_NOISE_PHILOX, MRMG_TYPE_HEUN
  public :: MRMG_SRC_NONE, MRMG_SRC_FIELD_TABLE, MRMG_SRC_TEMP_TABLE, MRMG_SRC_TEMP_SG
  public :: MRMG_FN_RAMP, MRMG_FN_SIN, MRMG_FN_EXP, MRMG_FN_COS, MRMG_FN_NEW, MRMG_FN_STE

  integer(c_int), parameter :: MRMG_FP64 = 0, MRMG_FP32 = 1
  integer(c_int), parameter :: MRMG_NOISE_NONE = 0, MRMG_NOISE_INJECT = 1, MRMG_NOISE_PHILOX = 2
  integer(c_int), parameter :: MRMG_TYPE_HEUN = -3
  ! moving sources of get_appl (applied.f90:827-1034) and the time functions of set_fld
  integer(c_int), parameter :: MRMG_SRC_NONE = 0, MRMG_SRC_FIELD_TABLE = 1, MRMG_SRC_TEMP_TABLE = 2, &
                               MRMG_SRC_TEMP_SG = 3
  integer(c_int), parameter :: MRMG_FN_RAMP = 0, MRMG_FN_SIN = 1, MRMG_FN_EXP = 2, MRMG_FN_COS = 3, &
                               MRMG_FN_NEW = 4, MRMG_FN_STE = 5

  type, bind(C) :: mrmg_geom
    integer(c_int) :: nx, ny, n_lay, precision, demag
    real(c_double) :: dx, dy
    type(c_ptr) :: dz, il_space
  end type

  type, bind(C) :: mrmg_layer
    integer(c_int) :: ng
    type(c_ptr) :: grid2grain, grain2gridindex, grain2grid
    type(c_ptr) :: grain_vol, ms, hk, uni_axis, cubic_axis
    type(c_ptr) :: hxgrainindex, hxgrain, hx
    type(c_ptr) :: uhxgrainindex, uhxgrain, uhx
    type(c_ptr) :: lhxgrainindex, lhxgrain, lhx
    type(c_ptr) :: ms_scale, alpha_scale, hk_scale, temp, h_app
    real(c_double) :: alpha, gamma, thm_alpha, floor_ms
    integer(c_int) :: thermal, use_mag
    type(c_ptr) :: tc, lambda_ms
    real(c_double) :: spin, tc_nom, msexp, hkexp
    integer(c_int) :: n_ms_tab, n_alpha_tab, n_hk_tab, pad3_     ! READ_SCALE tables
    type(c_ptr) :: ms_tab_t, ms_tab_s, alpha_tab_t, alpha_tab_s, hk_tab_t, hk_tab_s
  end type

  type, bind(C) :: mrmg_ufield
    real(c_double) :: dir(3)
    real(c_double) :: t_start, t_end, s_start, s_end, scale
    integer(c_int) :: func, pad_                ! MRMG_FN_*: 0 ramp, 1 SIN, 2 EXP, 3 COS, 4 NEW, 5 STE
    real(c_double) :: amp, gamma, phase, param, duty, moddepth, dc
  end type

  ! A moving field table (READ_FIELD data with FLD_VELOCITY), temperature table (READ_TEMP
  ! with TEMP_VELOCITY) or super-Gaussian spot of one layer: the field of get_appl's
  ! moving branches (applied.f90:884-903).  xpos(nxp), ypos(nyp) and vals (3,nxp,nyp) for a
  ! field or (nxp,nyp) for a temperature are passed with c_loc.
  type, bind(C) :: mrmg_source
    integer(c_int) :: kind, nxp, nyp, expn
    type(c_ptr) :: xpos, ypos, vals
    real(c_double) :: vel_x, vel_y, offset_x, offset_y, ref_x, ref_y, t_origin
    real(c_double) :: isigx, isigy, dtemp
    real(c_double) :: t_start, t_end, s_start, s_end, scale
    integer(c_int) :: func, pad_
    real(c_double) :: amp, gamma, phase, param, duty, moddepth, dc
  end type

  type, bind(C) :: mrmg_step_params
    integer(c_int) :: type, n_steps
    real(c_double) :: t0, dt
    integer(c_int) :: noise, pad_
    integer(c_int64_t) :: seed, step0
    real(c_double) :: dt_frac, dt_max, max_dm
    integer(c_int) :: llb, pad2_
  end type

  type, bind(C) :: mrmg_diag
    real(c_double) :: t, dt
    integer(c_int) :: steps, n_lay
    real(c_double) :: errmx, errmx1, mmag
    real(c_double) :: mave(3, 8)
    integer(c_int) :: rejected, pad_
    real(c_double) :: max_dmdt
    real(c_double) :: mvar(3, 8), have(3, 8)
    real(c_double) :: mave_all(3), mvar_all(3), have_all(3)
  end type

  ! Gaussian reciprocity sensitivity of a read head (w: c_loc of n_lay weights or c_null_ptr)
  type, bind(C) :: mrmg_sensor
    real(c_double) :: sigma_x, sigma_y
    real(c_double) :: dir(3)
    type(c_ptr) :: w
  end type

  interface
    integer(c_int) function mrmg_create(g, ctx) bind(C, name='mrmg_create')
      import :: c_int, c_ptr, mrmg_geom
      type(mrmg_geom), intent(in) :: g
      type(c_ptr), intent(out) :: ctx
    end function
    ! open in-plane boundaries: the demag over the media alone, on the 2nx x 2ny padded grid
    integer(c_int) function mrmg_create_open(g, ctx) bind(C, name='mrmg_create_open')
      import :: c_int, c_ptr, mrmg_geom
      type(mrmg_geom), intent(in) :: g
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mrmg_destroy(ctx) bind(C, name='mrmg_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! n_real: the (nx,ny,9,n_lay,n_lay) array of calculate_tensor (tensor.f90:82-216)
    integer(c_int) function mrmg_set_tensor(ctx, n_real) bind(C, name='mrmg_set_tensor')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: n_real(*)
    end function
    integer(c_int) function mrmg_build_tensor(ctx, num_images, n_real_out) bind(C, name='mrmg_build_tensor')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: num_images
      type(c_ptr), value :: n_real_out
    end function
    integer(c_int) function mrmg_build_tensor_images(ctx, num_images, sul_mu, head_mu, fly_height, space_to_sul, &
                                                     n_real_out) bind(C, name='mrmg_build_tensor_images')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: num_images
      real(c_double), value :: sul_mu, head_mu, fly_height, space_to_sul
      type(c_ptr), value :: n_real_out
    end function
    integer(c_int) function mrmg_set_layer(ctx, k, layer) bind(C, name='mrmg_set_layer')
      import :: c_int, c_ptr, mrmg_layer
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      type(mrmg_layer), intent(in) :: layer
    end function
    integer(c_int) function mrmg_set_ufield(ctx, k, uf) bind(C, name='mrmg_set_ufield')
      import :: c_int, c_ptr, mrmg_ufield
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      type(mrmg_ufield), intent(in) :: uf
    end function
    ! m: data_s%m(1+in:3+in,:) passed as a contiguous (3,ng) array
    integer(c_int) function mrmg_set_state(ctx, k, m) bind(C, name='mrmg_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double), intent(in) :: m(*)
    end function
    integer(c_int) function mrmg_get_state(ctx, k, m, h) bind(C, name='mrmg_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double), intent(out) :: m(*), h(*)
    end function
    ! mag: data_s%mag(in/3+1,:), the LLB |m| (ng)
    integer(c_int) function mrmg_set_mag(ctx, k, mag) bind(C, name='mrmg_set_mag')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double), intent(in) :: mag(*)
    end function
    integer(c_int) function mrmg_get_mag(ctx, k, mag) bind(C, name='mrmg_get_mag')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double), intent(out) :: mag(*)
    end function
    integer(c_int) function mrmg_step(ctx, p, noise, diag) bind(C, name='mrmg_step')
      import :: c_int, c_ptr, mrmg_step_params, mrmg_diag
      type(c_ptr), value :: ctx
      type(mrmg_step_params), intent(in) :: p
      type(c_ptr), value :: noise
      type(mrmg_diag), intent(out) :: diag
    end function
    integer(c_int) function mrmg_readback(ctx, sn, n_pos, xpos, ypos, out) bind(C, name='mrmg_readback')
      import :: c_int, c_ptr, c_double, mrmg_sensor
      type(c_ptr), value :: ctx
      type(mrmg_sensor), intent(in) :: sn
      integer(c_int), value :: n_pos
      real(c_double), intent(in) :: xpos(*), ypos(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mrmg_switch_arm(ctx, enable, dir, thr) bind(C, name='mrmg_switch_arm')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), intent(in) :: dir(3)
      real(c_double), value :: thr
    end function
    ! step: integer(c_int64_t) (ng); t (ng); h_app (3,ng)
    integer(c_int) function mrmg_switch_read(ctx, k, step, t, h_app) bind(C, name='mrmg_switch_read')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      integer(c_int64_t), intent(out) :: step(*)
      real(c_double), intent(out) :: t(*), h_app(*)
    end function
    ! slot 0: the field table of layer k; slot 1: its temperature table or spot
    integer(c_int) function mrmg_set_source(ctx, k, slot, src) bind(C, name='mrmg_set_source')
      import :: c_int, c_ptr, mrmg_source
      type(c_ptr), value :: ctx
      integer(c_int), value :: k, slot
      type(mrmg_source), intent(in) :: src
    end function
    ! launch shapes (the FFTW_EXHAUSTIVE planning role, fftw.f90:25); key: a C string,
    ! e.g. 'lanes_per_grain'//c_null_char
    integer(c_int) function mrmg_tune(ctx, key, value) bind(C, name='mrmg_tune')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      integer(c_int), value :: value
    end function
    integer(c_int) function mrmg_get_tune(ctx, key, value) bind(C, name='mrmg_get_tune')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      integer(c_int), intent(out) :: value
    end function
    ! the device's thermal normals for steps step0.., laid out (3,G,n_steps)
    integer(c_int) function mrmg_philox_normals(ctx, seed, step0, n_steps, out) bind(C, name='mrmg_philox_normals')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed, step0
      integer(c_int), value :: n_steps
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mrmg_footprint(ctx, bytes) bind(C, name='mrmg_footprint')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: bytes
    end function
    type(c_ptr) function mrmg_last_error() bind(C, name='mrmg_last_error')
      import :: c_ptr
    end function
  end interface

contains

  ! the last error message as a Fortran string
  function mrmg_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = mrmg_last_error()
    call c_f_pointer(p, s, [1024])
    n = 0
    do while (n < 1024)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module mrmg_m
