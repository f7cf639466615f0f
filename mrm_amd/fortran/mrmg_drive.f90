! mrmg_drive.f90 -- the smallest Fortran caller of the library, through mrmg_m.
!
! It does what the replacement llg() does for one LLG command: create a context, build
! the tensor on the device, hand over a layer in the reference's own array layouts
! (grid2grain(nx,ny), grain2grid(2,ncell), grain2gridindex(0:ng), hk(2,ng),
! uni_axis(3,ng), m(3,ng)), set a ramped uniform field, run Heun steps and read the
! state back.  The media is a 16 x 16 grid of 2 x 2-cell grains.  The final m is written
! to m_fortran.bin (raw FP64, Fortran order) in the working directory, so the same run
! made from another binding can be compared with it bit for bit.  A second run adds a
! moving write-head field table (the READ_FIELD data a FLD_VELOCITY command moves,
! applied.f90:884-903, through mrmg_set_source) and a launch shape set with mrmg_tune, and
! writes m_fortran_src.bin.
program mrmg_drive
  use mrmg_m
  use, intrinsic :: iso_c_binding
  implicit none
  integer, parameter :: nx = 16, ny = 16, ng = 64, nsteps = 50
  real(c_double), parameter :: dx = 2d-7, dy = 2d-7
  type(mrmg_geom) :: g
  type(mrmg_layer) :: lay
  type(mrmg_ufield) :: uf
  type(mrmg_source) :: src
  type(mrmg_step_params) :: sp
  type(mrmg_diag) :: dg
  type(c_ptr) :: ctx
  real(c_double), target :: dz(1), il(1)
  integer(c_int), target :: g2g(nx, ny), g2gi(0:ng), cells(2, nx * ny)
  real(c_double), target :: vol(ng), ms(ng), hk(2, ng), uni(3, ng), m(3, ng), h(3, ng)
  real(c_double) :: r2
  integer :: i, j, n, q, run
  integer, parameter :: nxp = 5, nyp = 4
  real(c_double), target :: xp(nxp), yp(nyp), tab(3, nxp, nyp), m0(3, ng)
  integer(c_int) :: lanes

  dz(1) = 4d-7
  il(1) = 0d0
  do j = 1, ny
    do i = 1, nx
      g2g(i, j) = ((j - 1) / 2) * (nx / 2) + (i - 1) / 2 + 1
    end do
  end do
  q = 0
  g2gi(0) = 0
  do n = 1, ng                       ! each grain's cells, x fastest
    do j = 1, ny
      do i = 1, nx
        if (g2g(i, j) == n) then
          q = q + 1
          cells(1, q) = i
          cells(2, q) = j
        end if
      end do
    end do
    g2gi(n) = q
  end do
  do n = 1, ng
    vol(n) = 4d0 * dx * dy * dz(1)
    ms(n) = 700d0
    hk(1, n) = 1d4 + 100d0 * n
    hk(2, n) = 0d0
    ! easy axes tilted a little, built without transcendental functions so that another
    ! binding can reproduce them bit for bit (sqrt is correctly rounded everywhere)
    uni(:, n) = (/ 0.01d0 * mod(n, 5), 0.02d0 * mod(n, 3), 1d0 /)
    r2 = uni(1, n) * uni(1, n) + uni(2, n) * uni(2, n) + uni(3, n) * uni(3, n)
    uni(:, n) = uni(:, n) / sqrt(r2)
    m(:, n) = (/ 0.1d0, 0d0, 0.995d0 /)
    r2 = m(1, n) * m(1, n) + m(2, n) * m(2, n) + m(3, n) * m(3, n)
    m(:, n) = m(:, n) / sqrt(r2)
  end do

  m0 = m
  ! a head-field table on a coarse mesh, polynomial in the node position (no transcendental
  ! functions, so another binding builds the same bits)
  do i = 1, nxp
    xp(i) = -2d-6 + 1.2d-6 * (i - 1)
  end do
  do j = 1, nyp
    yp(j) = -1d-6 + 1.5d-6 * (j - 1)
  end do
  do j = 1, nyp
    do i = 1, nxp
      tab(1, i, j) = 300d0 * (i - 3)
      tab(2, i, j) = 50d0 * (j - 2) * (i - 1)
      tab(3, i, j) = -2d3 - 500d0 * (i - 1) + 100d0 * (j - 1) * (j - 1)
    end do
  end do

  do run = 1, 2
    g = mrmg_geom(nx, ny, 1, MRMG_FP64, 1, dx, dy, c_loc(dz), c_loc(il))
    call check(mrmg_create(g, ctx), 'mrmg_create')
    call check(mrmg_build_tensor(ctx, 1, c_null_ptr), 'mrmg_build_tensor')

    lay%ng = ng
    lay%grid2grain = c_loc(g2g);  lay%grain2gridindex = c_loc(g2gi);  lay%grain2grid = c_loc(cells)
    lay%grain_vol = c_loc(vol);  lay%ms = c_loc(ms);  lay%hk = c_loc(hk);  lay%uni_axis = c_loc(uni)
    lay%cubic_axis = c_null_ptr
    lay%hxgrainindex = c_null_ptr;  lay%hxgrain = c_null_ptr;  lay%hx = c_null_ptr
    lay%uhxgrainindex = c_null_ptr;  lay%uhxgrain = c_null_ptr;  lay%uhx = c_null_ptr
    lay%lhxgrainindex = c_null_ptr;  lay%lhxgrain = c_null_ptr;  lay%lhx = c_null_ptr
    lay%ms_scale = c_null_ptr;  lay%alpha_scale = c_null_ptr;  lay%hk_scale = c_null_ptr
    lay%temp = c_null_ptr;  lay%h_app = c_null_ptr;  lay%tc = c_null_ptr;  lay%lambda_ms = c_null_ptr
    lay%alpha = 0.1d0;  lay%gamma = 1.76d7;  lay%thm_alpha = 1d0;  lay%floor_ms = 0.01d0
    lay%thermal = 0;  lay%use_mag = 0
    lay%spin = 0.5d0;  lay%tc_nom = 1d22;  lay%msexp = -1d0;  lay%hkexp = -1d0
    lay%n_ms_tab = 0;  lay%n_alpha_tab = 0;  lay%n_hk_tab = 0;  lay%pad3_ = 0
    lay%ms_tab_t = c_null_ptr;  lay%ms_tab_s = c_null_ptr;  lay%alpha_tab_t = c_null_ptr
    lay%alpha_tab_s = c_null_ptr;  lay%hk_tab_t = c_null_ptr;  lay%hk_tab_s = c_null_ptr
    call check(mrmg_set_layer(ctx, 0, lay), 'mrmg_set_layer')

    uf%dir = (/ 0.1d0, 0d0, -1d0 /) / sqrt(1.01d0)
    uf%t_start = 0d0;  uf%t_end = 1d-10;  uf%s_start = 0d0;  uf%s_end = 1d4;  uf%scale = 1d0
    uf%func = MRMG_FN_RAMP;  uf%pad_ = 0
    uf%amp = 0d0;  uf%gamma = 0d0;  uf%phase = 0d0;  uf%param = 0d0;  uf%duty = 0d0
    uf%moddepth = 0d0;  uf%dc = 0d0
    call check(mrmg_set_ufield(ctx, 0, uf), 'mrmg_set_ufield')
    if (run == 2) then
      ! the table moves down-track at 2000 cm/s and its field ramps in over 2e-12 s
      src%kind = MRMG_SRC_FIELD_TABLE;  src%nxp = nxp;  src%nyp = nyp;  src%expn = 2
      src%xpos = c_loc(xp);  src%ypos = c_loc(yp);  src%vals = c_loc(tab)
      src%vel_x = 0d0;  src%vel_y = 2000d0;  src%offset_x = 0d0;  src%offset_y = 0d0
      src%ref_x = 0d0;  src%ref_y = 0d0;  src%t_origin = 0d0
      src%isigx = 0d0;  src%isigy = 0d0;  src%dtemp = 0d0
      src%t_start = 0d0;  src%t_end = 2d-12;  src%s_start = 0d0;  src%s_end = 1d0;  src%scale = 1d0
      src%func = MRMG_FN_RAMP;  src%pad_ = 0
      src%amp = 0d0;  src%gamma = 0d0;  src%phase = 0d0;  src%param = 0d0;  src%duty = 0d0
      src%moddepth = 0d0;  src%dc = 0d0
      call check(mrmg_set_source(ctx, 0, 0, src), 'mrmg_set_source')
      call check(mrmg_tune(ctx, 'lanes_per_grain'//c_null_char, 4_c_int), 'mrmg_tune')
      call check(mrmg_get_tune(ctx, 'lanes_per_grain'//c_null_char, lanes), 'mrmg_get_tune')
      if (lanes /= 4) then
        print '(a)', '!--- mrmg_get_tune returned the wrong value'
        stop 1
      end if
    end if
    m = m0
    call check(mrmg_set_state(ctx, 0, m), 'mrmg_set_state')

    sp%type = MRMG_TYPE_HEUN;  sp%n_steps = nsteps;  sp%t0 = 0d0;  sp%dt = 5d-14
    sp%noise = MRMG_NOISE_NONE;  sp%pad_ = 0;  sp%seed = 0;  sp%step0 = 0
    sp%dt_frac = 0d0;  sp%dt_max = 1d22;  sp%max_dm = 180d0;  sp%llb = 0;  sp%pad2_ = 0
    call check(mrmg_step(ctx, sp, c_null_ptr, dg), 'mrmg_step')
    call check(mrmg_get_state(ctx, 0, m, h), 'mrmg_get_state')
    call check(mrmg_destroy(ctx), 'mrmg_destroy')

    if (run == 1) then
      open(10, file='m_fortran.bin', access='stream', form='unformatted', status='replace')
    else
      open(10, file='m_fortran_src.bin', access='stream', form='unformatted', status='replace')
    end if
    write(10) m
    close(10)
    print '(a,i0,a,i0,a,es24.16,a,es24.16)', 'mrmg_drive run ', run, ': ', dg%steps, ' steps, t = ', dg%t, &
      ', <m_z> = ', dg%mave(3, 1)
  end do

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= 0) then
      print '(a)', '!--- '//what//' failed: '//mrmg_error()
      stop 1
    end if
  end subroutine

end program
