! hydra_driver.f90 -- drop-in driver for the HydraPhase step loop on one
! MI355X.  It reads the reference's own inputs (flow.in, itter.in,
! geometry_files/grid_proc0000.dat, bc_proc0000.in, proc_bc0000.in,
! restart_files/restart0000.in) and runs the reference's step sequence
! (main.F90:511-657) with every routine executed by libhydra_hip.so through
! hydra_hip_mod.  Per step it prints the reference's monitor line
! (main.F90:653) and appends time and |u_g| of cell 1 to Velocity.dat
! (:641,654); at the end it writes restart_files/restart0000.in (:812-828)
! and tecplot_files/tecplot0000.dat (:956-1266).
!
! HYDRA_FUSED=1 in the environment runs each step as the fused,
! device-resident hp_step_diffs call (the time step reduced and applied on
! the device, the monitor reduced there), then reads back the monitor, the
! time and cell 1's primitives for the same two outputs.
program hydra_driver
  use iso_c_binding
  use hydra_hip_mod
  implicit none
  integer, parameter :: dp = selected_real_kind(15, 307)
  type(hp_params) :: prm
  type(c_ptr) :: ctx
  integer :: nodes, neles, nghosts, pghosts, max_pghosts, ntot, i, k, n, nel, itter, total
  integer :: iter, iters, niter, itersub, istart, myflux, myorder, u, ios, dummy, g_id
  real(dp), allocatable :: x(:), y(:), z(:), cu(:, :), gp(:, :)
  integer, allocatable :: nod(:, :), nc(:, :), iblank(:), nparent(:), nghost(:), ntype(:), nside(:)
  real(dp) :: time, dt, dtmin, diffs(4), pr(14), vel_mag, scal(3)
  real(dp) :: rogmin, ropmin, egtotalmin, eptotalmin, phig0, phip0
  character(len=16) :: fused_env
  logical :: fused

  call read_flow(prm, niter, itersub, istart, myflux, myorder)
  prm%precision = 0
  call get_environment_variable('HYDRA_FUSED', fused_env)
  fused = (trim(fused_env) == '1')

  ! mesh (sr_geometry.F90:101-139)
  open(newunit=u, file='geometry_files/grid_proc0000.dat', status='old')
  read(u, *) nodes, neles
  allocate(x(nodes), y(nodes), z(nodes))
  do i = 1, nodes
    read(u, *) n, x(n), y(n), z(n), g_id
  end do
  allocate(nod(neles, 4), nc(neles, 4), iblank(neles))
  do i = 1, neles
    read(u, *) nel, nod(nel, 1), nod(nel, 2), nod(nel, 3), nod(nel, 4), &
               nc(nel, 1), nc(nel, 2), nc(nel, 3), nc(nel, 4), iblank(nel)
  end do
  close(u)
  open(newunit=u, file='geometry_files/bc_proc0000.in', status='old')
  read(u, *) nghosts
  allocate(nparent(nghosts), nghost(nghosts), ntype(nghosts), nside(nghosts))
  do i = 1, nghosts
    read(u, *) nparent(i), nghost(i), ntype(i), nside(i)
  end do
  close(u)
  open(newunit=u, file='geometry_files/proc_bc0000.in', status='old')
  read(u, *) pghosts, max_pghosts
  close(u)
  if (pghosts /= 0) then
    print *, 'hydra_driver runs one partition; proc_bc0000.in lists ', pghosts, ' partition ghosts'
    stop 1
  end if
  ntot = neles + nghosts

  allocate(cu(neles, 10), gp(nghosts, 6))
  total = 0
  itter = 0
  if (istart > 0) then
    ! main.F90:427-441
    open(newunit=u, file='restart_files/restart0000.in', status='old', action='read')
    read(u, *) time
    read(u, *) ((cu(i, k), i = 1, neles), k = 1, 10)
    do k = 1, 6
      read(u, *) (gp(i, k), i = 1, nghosts)
    end do
    close(u)
    open(newunit=u, file='itter.in', status='old')
    read(u, *, iostat=ios) itter
    close(u)
  else
    ! uniform free stream, sr_init_cond.F90:177-195 with main.F90:391-402
    time = 0.0_dp
    rogmin = prm%pmin / (287.0_dp * prm%tgmin)
    ropmin = (prm%pmin + prm%pinf) * prm%gamap / ((prm%gamap - 1.0_dp) * prm%cpp * prm%tpmin)
    egtotalmin = 287.0_dp * prm%tgmin / (prm%gamag - 1.0_dp) + 0.5_dp * (prm%ugmin**2 + prm%vgmin**2 + prm%wgmin**2)
    eptotalmin = (prm%cpp * prm%tpmin) / prm%gamap + prm%pinf / ropmin + 0.5_dp * (prm%upmin**2 + prm%vpmin**2 + prm%wpmin**2)
    phip0 = prm%epslninf
    phig0 = 1 - prm%epslninf
    do i = 1, neles
      cu(i, 1) = phig0 * rogmin
      cu(i, 2) = phig0 * rogmin * prm%ugmin
      cu(i, 3) = phig0 * rogmin * prm%vgmin
      cu(i, 4) = phig0 * rogmin * prm%wgmin
      cu(i, 5) = phig0 * rogmin * egtotalmin
      cu(i, 6) = phip0 * ropmin
      cu(i, 7) = phip0 * ropmin * prm%upmin
      cu(i, 8) = phip0 * ropmin * prm%vpmin
      cu(i, 9) = phip0 * ropmin * prm%wpmin
      cu(i, 10) = phip0 * ropmin * eptotalmin
    end do
    ! ghost gas primitives start at zero (ALLOCATE leaves them unset in
    ! the reference; the first BC applications overwrite them)
    gp = 0.0_dp
  end if

  call hp_check(hp_create(ctx, nodes, neles, nghosts, x, y, z, nod, neles, nc(:, 1), nc(:, 2), nc(:, 3), &
                          nc(:, 4), iblank, nparent, nghost, ntype, nside, prm), c_null_ptr, 'hp_create')
  call hp_check(hp_upload_state(ctx, cu, neles, gp, time, total), ctx, 'hp_upload_state')

  open(unit=23, file='Velocity.dat', access='append')
  do iter = 1, niter
    if (fused) then
      do iters = 1, itersub
        call hp_check(hp_step_diffs(ctx, 1, diffs), ctx, 'hp_step_diffs')
        total = total + 1
        call hp_check(hp_get_array(ctx, 11, scal, 3_c_size_t), ctx, 'scalars')  ! dt, time, total
        time = scal(2)
        call hp_check(hp_probe_cell(ctx, 1, pr), ctx, 'probe')
        vel_mag = sqrt(pr(1)**2 + pr(2)**2 + pr(3)**2)
        print '(I6, 1PE20.10, 5F24.16)', iters, time, diffs(1), diffs(2), diffs(3), diffs(4)
        write(23, *) time, vel_mag
      end do
      cycle
    end if
    do iters = 1, itersub
      ! main.F90:515-558
      call hp_check(hp_primitive(ctx), ctx, 'primitive')
      total = total + 1
      call hp_check(hp_dt_min(ctx, dtmin), ctx, 'dt')
      dt = dtmin * prm%cfl
      time = time + dt
      call hp_check(hp_set_dt(ctx, dt, time, total), ctx, 'set_dt')
      call hp_check(hp_smooth(ctx), ctx, 'smooth')          ! :563
      call hp_check(hp_rk(ctx), ctx, 'rk')                   ! :578
      call hp_check(hp_end_step(ctx, diffs), ctx, 'end_step') ! :595-625
      call hp_check(hp_probe_cell(ctx, 1, pr), ctx, 'probe')
      vel_mag = sqrt(pr(1)**2 + pr(2)**2 + pr(3)**2)
      print '(I6, 1PE20.10, 5F24.16)', iters, time, diffs(1), diffs(2), diffs(3), diffs(4)
      write(23, *) time, vel_mag
    end do
  end do
  close(23)

  call hp_check(hp_download_state(ctx, cu, neles, gp, time, total), ctx, 'hp_download_state')
  call hp_check(hp_destroy(ctx), c_null_ptr, 'hp_destroy')
  ! main.F90:812-828 (17 significant digits so the values round-trip)
  call execute_command_line('mkdir -p restart_files')
  open(newunit=u, file='restart_files/restart0000.in', status='replace', action='write')
  write(u, '(1X,ES24.16E3)') time
  write(u, '(*(1X,ES24.16E3))') ((cu(i, k), i = 1, neles), k = 1, 10)
  do k = 1, 6
    write(u, '(*(1X,ES24.16E3))') (gp(i, k), i = 1, nghosts)
  end do
  close(u)
  open(newunit=u, file='itter.in', status='replace')
  write(u, *) itter
  close(u)
  call write_tecplot()

contains

  ! main.F90:956-1266: Tecplot FE file of the final state, primitives from
  ! cu by the reference's two-phase recovery; the reference names 16
  ! variables and then writes 17 blocks (PHIP last), reproduced as is
  subroutine write_tecplot()
    real(dp), allocatable :: w(:, :)
    real(dp) :: eg, ep, bbb, ccc, psig, fg, rrg, s4
    integer :: ie, uo, j
    allocate(w(neles, 14))
    rrg = 287.0_dp
    s4 = 4.0_dp * prm%surfacetension / prm%diap
    do ie = 1, neles
      associate (ug => w(ie, 1), vg => w(ie, 2), wg => w(ie, 3), pg => w(ie, 4), pp => w(ie, 5), &
                 tg => w(ie, 6), up => w(ie, 7), vp => w(ie, 8), wp => w(ie, 9), phig => w(ie, 10), &
                 tp => w(ie, 11), rog => w(ie, 12), rop => w(ie, 13), phip => w(ie, 14))
        ug = cu(ie, 2) / cu(ie, 1); vg = cu(ie, 3) / cu(ie, 1); wg = cu(ie, 4) / cu(ie, 1)
        up = cu(ie, 7) / cu(ie, 6); vp = cu(ie, 8) / cu(ie, 6); wp = cu(ie, 9) / cu(ie, 6)
        eg = cu(ie, 5) / cu(ie, 1) - 0.5_dp * (ug**2 + vg**2 + wg**2)
        ep = cu(ie, 10) / cu(ie, 6) - 0.5_dp * (up**2 + vp**2 + wp**2)
        bbb = (s4 + prm%pinf * prm%gamap) - (cu(ie, 1) * (prm%gamag - 1.0_dp) * eg) - &
              (cu(ie, 6) * (prm%gamap - 1.0_dp) * ep)
        ccc = (cu(ie, 1) * (prm%gamag - 1.0_dp) * eg * (s4 + prm%pinf * prm%gamap))
        pg = (-bbb + sqrt(bbb**2 + 4 * ccc)) * 0.5_dp
        pp = pg + s4
        phig = (cu(ie, 1) * (prm%gamag - 1.0_dp) * eg) / pg
        phip = 1.0_dp - phig
        rog = cu(ie, 1) / phig
        rop = cu(ie, 6) / phip
        tg = pg / (rog * rrg)
        tp = ((pp + prm%pinf) * prm%gamap) / ((prm%gamap - 1.0_dp) * rop * prm%cpp)
        if (phig < prm%epslnmin) then
          phig = prm%epslnmin; phip = 1.0_dp - phig
          ug = up; vg = vp; wg = wp; tg = tp; pg = pp
        end if
        if (phip < prm%epslnmin) then
          phip = prm%epslnmin; phig = 1.0_dp - phip
          up = ug; vp = vg; wp = wg; tp = tg; pp = pg
        end if
        if (phig >= prm%epslnmin .and. phig <= prm%epslnmax) then
          psig = (phig - prm%epslnmin) / (prm%epslnmax - prm%epslnmin)
          fg = -psig * psig * (2.0_dp * psig - 3.0_dp)
          ug = fg * ug + (1.0_dp - fg) * up; vg = fg * vg + (1.0_dp - fg) * vp
          wg = fg * wg + (1.0_dp - fg) * wp; tg = fg * tg + (1.0_dp - fg) * tp
          pg = fg * pg + (1.0_dp - fg) * pp
        end if
        if (phip >= prm%epslnmin .and. phip <= prm%epslnmax) then
          psig = (phip - prm%epslnmin) / (prm%epslnmax - prm%epslnmin)
          fg = -psig * psig * (2.0_dp * psig - 3.0_dp)
          up = fg * up + (1.0_dp - fg) * ug; vp = fg * vp + (1.0_dp - fg) * vg
          wp = fg * wp + (1.0_dp - fg) * wg; tp = fg * tp + (1.0_dp - fg) * tg
          pp = fg * pp + (1.0_dp - fg) * pg
        end if
        rog = pg / (rrg * tg)
        rop = (pp + prm%pinf) * prm%gamap / ((prm%gamap - 1.0_dp) * prm%cpp * tp)
      end associate
    end do
    call execute_command_line('mkdir -p tecplot_files')
    open(newunit=uo, file='tecplot_files/tecplot0000.dat', status='replace', action='write')
    write(uo, '(A)') 'VARIABLES = "X", "Y", "Z", "U","V", "W", "PG", "PP", "TG", ' // &
                     '"UP", "VP", "WP", "PHIG", "TP", "ROG", "ROP"'
    write(uo, '(A,I0,A,I0,A,I0,A)') 'ZONE ZONETYPE=FETETRAHEDRON, NODES=', nodes, ', ELEMENTS=', neles, &
        ', DATAPACKING=BLOCK, VARLOCATION=([4-', 16, ']=CELLCENTERED)'
    write(uo, '(ES16.8)') x
    write(uo, '(ES16.8)') y
    write(uo, '(ES16.8)') z
    do j = 1, 14
      write(uo, '(ES16.8)') w(:, j)
    end do
    do ie = 1, neles
      write(uo, *) nod(ie, 1), nod(ie, 2), nod(ie, 3), nod(ie, 4)
    end do
    close(uo)
  end subroutine write_tecplot

  ! sr_Read_flow.F90:19-268: each tuple is the next line that is neither
  ! blank nor a '#' comment and parses, with any trailing comment removed
  subroutine next_line(un, line, ok)
    integer, intent(in) :: un
    character(len=256), intent(out) :: line
    logical, intent(out) :: ok
    integer :: st, cp
    ok = .false.
    do
      read(un, '(A)', iostat=st) line
      if (st /= 0) return
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '#') cycle
      cp = index(line, '#')
      if (cp > 0) line = line(1:cp - 1)
      ok = .true.
      return
    end do
  end subroutine next_line

  ! the next tuple of n values: lines that do not parse are skipped as the
  ! reference's per-tuple loops do
  subroutine next_tuple(un, n, v)
    integer, intent(in) :: un, n
    real(dp), intent(out) :: v(n)
    character(len=256) :: line
    logical :: ok
    integer :: st
    v = 0.0_dp
    do
      call next_line(un, line, ok)
      if (.not. ok) return
      read(line, *, iostat=st) v
      if (st == 0) return
    end do
  end subroutine next_tuple

  subroutine read_flow(p, niter_, itersub_, istart_, myflux_, myorder_)
    type(hp_params), intent(out) :: p
    integer, intent(out) :: niter_, itersub_, istart_, myflux_, myorder_
    integer :: un, st
    real(dp) :: v(3)
    p%precision = 0
    p%co2 = 0; p%co4 = 0; p%c23 = 0; p%cmug = 0; p%rrg = 0; p%gravity = 0; p%prp = 0; p%foursigmabyd = 0
    open(newunit=un, file='flow.in', status='old', action='read', iostat=st)
    if (st /= 0) then
      print *, 'Error opening flow.in, IOSTAT=', st
      stop 1
    end if
    call next_tuple(un, 2, v); p%pmax = v(1); p%pmin = v(2)
    call next_tuple(un, 2, v); p%tgmax = v(1); p%tgmin = v(2)
    call next_tuple(un, 2, v); p%ugmax = v(1); p%ugmin = v(2)
    call next_tuple(un, 2, v); p%vgmax = v(1); p%vgmin = v(2)
    call next_tuple(un, 2, v); p%wgmax = v(1); p%wgmin = v(2)
    call next_tuple(un, 2, v); p%upmax = v(1); p%upmin = v(2)
    call next_tuple(un, 2, v); p%vpmax = v(1); p%vpmin = v(2)
    call next_tuple(un, 2, v); p%wpmax = v(1); p%wpmin = v(2)
    call next_tuple(un, 2, v); p%tpmax = v(1); p%tpmin = v(2)
    call next_tuple(un, 2, v); p%pinf = v(1); p%gamap = v(2)
    call next_tuple(un, 2, v); p%gamag = v(1); p%cpp = v(2)
    call next_tuple(un, 2, v); p%cpg = v(1); p%amolg = v(2)
    call next_tuple(un, 2, v); p%prg = v(1); p%epslnmax = v(2)
    call next_tuple(un, 3, v); p%epslnmin = v(1); p%epslninf = v(2); p%phipzero = v(3)
    call next_tuple(un, 3, v); p%cfl = v(1); niter_ = nint(v(2)); itersub_ = nint(v(3))
    call next_tuple(un, 2, v); istart_ = nint(v(1)); p%diap = v(2)
    call next_tuple(un, 2, v); p%surfacetension = v(1); myflux_ = nint(v(2))
    call next_tuple(un, 3, v); myorder_ = nint(v(1)); p%as = v(2); p%ts = v(3)
    close(un)
    p%niter = niter_; p%itersub = itersub_; p%istart = istart_
    p%myflux = myflux_; p%myorder = myorder_
  end subroutine read_flow

end program hydra_driver
