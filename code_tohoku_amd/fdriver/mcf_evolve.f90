! mcf_evolve.f90 -- Fortran driver: the reference's EVOLVE loop on one GPU.
!
! Follows 3DMHD-CfCA/MCF3DSGincfca3.f90 ("M3"): INPUT3D.DAT (M3:134-151),
! the cell-centre file cdnt.DAT (M3:300-307), tsave.DAT / tsave2D.DAT /
! time.DAT (M3:309-312, 520-530), the restart files 000{rank}.dat of an
! NPE-rank run (M3:454-465, 749-757), the EVOLVE outer/inner loops with the
! tsave/tfinal stops, the 2D-slice cadence and the dt print (M3:554-670),
! the wall-clock guard that ends a run after 11.7 hours (M3:663-670), SAVEU's
! FP32 snapshot, FP64 restart, time.DAT, tsave.DAT and info{nunit}.DAT
! (M3:684-757) and SAVEU2D's k = 1 slices 2D{nunit2D}{rank}.dat with
! tsave2D.DAT (M3:763-794).  The whole domain lives on the GPU; each physics
! call of the inner loop is one call through the C ABI (mcf_timestep for
! Couran+Stblty+reduction, mcf_step for SOURCE / MHD x3 / CT / SOURCE /
! DISSIP, with GRAVTY's kicks and solve when ifgrv = 2; the restart and
! snapshot files then carry the potential Phi).
!
! Environment: MCF_WALL_HOURS overrides the 11.7-hour wall-clock limit;
! MCF_COOLING=chem selects the reference's chemistry-network Fcool (default:
! the Koyama-Inutsuka Fcool); MCF_PRECISION=mixed the mixed-precision path.
!
! usage: mcf_evolve [rundir [NPE]]   (NPE: rank count of the reproduced run,
!                                     default 8; fixes the block split)
program mcf_evolve
  use iso_c_binding
  use mcf_abi_mod
  implicit none

  type(mcf_config) :: cfg
  type(c_ptr) :: ctx
  real(c_double), allocatable, target :: U(:,:,:,:), SP(:,:,:,:), BC(:,:,:,:), PH(:,:,:)
  character(len=512) :: rundir, arg
  integer :: npe, nsx, nsy, nsz, nx, ny, nz, nunit, nunit2D, st, itime, in10, in20, ifevo
  integer(c_int) :: rc, tag
  real(c_double) :: time, dt, tsave, tfinal, tsave2D
  real(c_double), parameter :: dtsave = 0.25d0, dtsave2D = 0.025d0
  real(c_double) :: t(1000), time_cpu(3), wall_limit
  integer(8) :: clk0, clk1, clkrate
  integer :: i, ios

  rundir = '.'
  npe = 8
  if (command_argument_count() >= 1) call get_command_argument(1, rundir)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) npe
  end if
  call nsplit_table(npe, nsx, nsy, nsz)

  call mcf_default_config(cfg)
  call read_input3d(trim(rundir)//'/INPUT3D.DAT', cfg)
  cfg%nsplitx = nsx; cfg%nsplity = nsy; cfg%nsplitz = nsz
  nx = cfg%np1x + cfg%np2x; ny = cfg%np1y + cfg%np2y; nz = cfg%np1z + cfg%np2z
  tfinal = cfg%tfinal
  call write_cdnt(trim(rundir), cfg, nx, ny, nz)
  wall_limit = 11.7d0
  call get_environment_variable('MCF_WALL_HOURS', arg, status=ios)
  if (ios == 0) read(arg, *) wall_limit
  call get_environment_variable('MCF_COOLING', arg, status=ios)
  if (ios == 0) then
    select case (trim(arg))
    case ('chem')
      cfg%cooling = MCF_COOL_CHEM
    case ('ki')
      cfg%cooling = MCF_COOL_KI
    case default
      write(0, *) 'MCF_COOLING must be ki or chem'
      stop 2
    end select
  end if
  call get_environment_variable('MCF_PRECISION', arg, status=ios)
  if (ios == 0) then
    select case (trim(arg))
    case ('mixed')
      cfg%precision = MCF_PREC_MIXED
    case ('fp64')
      cfg%precision = MCF_PREC_FP64
    case default
      write(0, *) 'MCF_PRECISION must be fp64 or mixed'
      stop 2
    end select
  end if

  open(2, file=trim(rundir)//'/tsave.DAT')
  read(2, *) time
  read(2, *) nunit
  close(2)
  nunit2D = 0
  open(2, file=trim(rundir)//'/tsave2D.DAT', status='old', iostat=ios)
  if (ios == 0) then
    read(2, *) nunit2D
    close(2)
  end if
  t = 0.d0
  open(3, file=trim(rundir)//'/time.DAT')
  do i = 1, nunit
    read(3, '(1p1d25.17)', end=10) t(i)
  end do
10 close(3)

  allocate(U(-1:nx+2, -1:ny+2, -1:nz+2, 8), SP(-1:nx+2, -1:ny+2, -1:nz+2, 8), BC(-1:nx+2, -1:ny+2, -1:nz+2, 3))
  allocate(PH(-1:nx+2, -1:ny+2, -1:nz+2))
  PH = 0.d0
  U = 0.d0; SP = 0.d0
  if (nunit == 1) then
    call shock_tube_ic(cfg, U, SP, nx, ny, nz)
  else
    call read_restart(trim(rundir), U, SP, nx, ny, nz, nsx, nsy, nsz)
  end if

  call mcf_check(mcf_init(cfg, ctx), 'mcf_init')
  call mcf_check(mcf_upload(ctx, U, SP), 'mcf_upload')
  call mcf_check(mcf_set_time(ctx, time, 1_c_int), 'mcf_set_time')

  open(5, file=trim(rundir)//'/info'//i3(nunit)//'.DAT')
  st = 1
  dt = 0.d0
  itime = 1 + int((time + 1.d-8)/dtsave)
  time_cpu = 0.d0
  call system_clock(count_rate=clkrate)
  do in10 = 1, cfg%maxstp
    call system_clock(clk0)
    tsave = dtsave*dble(itime)
    if (time >= tfinal) goto 9000
    if (time >= tsave) goto 7777
    call saveu(0)
    do in20 = 1, cfg%nitera
      tsave2D = dtsave2D*nunit2D
      if (time >= tsave2D) call saveu2d()
      if (time >= tfinal) goto 9000
      if (time >= tsave) goto 7777
      call mcf_check(mcf_timestep(ctx, tsave, dt, tag), 'mcf_timestep')
      st = tag
      write(*, *) in20, time, dt
      call mcf_check(mcf_step(ctx, dt), 'mcf_step')
      call mcf_check(mcf_get_time(ctx, time, ifevo), 'mcf_get_time')
    end do
    itime = itime - 1
7777 continue
    itime = itime + 1
    ! wall-clock guard (M3:660-670): hours of this outer iteration, summed
    call system_clock(clk1)
    time_cpu(2) = dble(clk1 - clk0)/dble(clkrate)/3.6d3
    time_cpu(3) = time_cpu(3) + time_cpu(2)
    if (time_cpu(3) + time_cpu(2) > wall_limit) goto 9000
  end do
9000 continue
  call saveu(1)
  call mcf_check(mcf_finalize(ctx), 'mcf_finalize')

contains

  function i3(k) result(s)
    integer, intent(in) :: k
    character(len=3) :: s
    write(s, '(I3.3)') k
  end function

  ! SAVEU (M3:684-757): FP32 snapshot, FP64 restart, counters, info line
  subroutine saveu(msig)
    integer, intent(in) :: msig
    character(len=7), parameter :: stb(2) = ['FastSpd', 'Conduct']
    call mcf_check(mcf_download(ctx, U, SP), 'mcf_download')
    if (cfg%ifgrv == 2) call mcf_check(mcf_download_phi(ctx, PH), 'mcf_download_phi')
    call write_rank_files(trim(rundir), U, SP, PH, nx, ny, nz, nsx, nsy, nsz, nunit)
    t(nunit) = time
    open(3, file=trim(rundir)//'/time.DAT')
    do i = 1, nunit
      write(3, '(1p1d25.17)') t(i)
    end do
    close(3)
    write(5, '(1p1d25.17,a,i8,a,1p1e11.3,a)') time, '  Step =', (nunit-1)*cfg%nitera, '  dt =', dt, stb(st)
    nunit = nunit + 1
    write(5, '(a,1p1e11.3,1p1e11.3)') 'Done ! Time =', time, tfinal
    open(2, file=trim(rundir)//'/tsave.DAT')
    write(2, '(1p1d25.17)') time
    write(2, '(i8)') nunit - 1
    close(2)
    if (msig == 1) close(5)
  end subroutine saveu

  ! SAVEU2D (M3:763-794): each rank's k = 1 plane as FP32 records, Bcc in
  ! place of the face field, then the slice counter and time
  subroutine saveu2d()
    call mcf_check(mcf_download(ctx, U, SP), 'mcf_download')
    call mcf_check(mcf_download_bcc(ctx, BC), 'mcf_download_bcc')
    if (cfg%ifgrv == 2) call mcf_check(mcf_download_phi(ctx, PH), 'mcf_download_phi')
    call write_rank_slices(trim(rundir), U, SP, BC, PH, nx, ny, nz, nsx, nsy, nsz, nunit2D)
    nunit2D = nunit2D + 1
    open(3, file=trim(rundir)//'/tsave2D.DAT')
    write(3, *) nunit2D
    write(3, *) time
    close(3)
  end subroutine saveu2d

end program mcf_evolve


! rank layout of the reference (M3:64-72)
subroutine nsplit_table(npe, sx, sy, sz)
  integer, intent(in) :: npe
  integer, intent(out) :: sx, sy, sz
  select case (npe)
  case (1);    sx = 1; sy = 1; sz = 1
  case (4);    sx = 2; sy = 2; sz = 1
  case (8);    sx = 2; sy = 2; sz = 2
  case (16);   sx = 4; sy = 2; sz = 2
  case (32);   sx = 4; sy = 4; sz = 2
  case (64);   sx = 4; sy = 4; sz = 4
  case (128);  sx = 8; sy = 4; sz = 4
  case (256);  sx = 8; sy = 8; sz = 4
  case (512);  sx = 8; sy = 8; sz = 8
  case (1024); sx = 8; sy = 8; sz = 16
  case default
    write(0, *) 'NPE not in the reference table:', npe
    stop 2
  end select
end subroutine nsplit_table


! INPUT3D.DAT, read the way INITIA reads it (M3:134-151)
subroutine read_input3d(path, cfg)
  use mcf_abi_mod
  character(*), intent(in) :: path
  type(mcf_config), intent(inout) :: cfg
  integer :: np1x, np2x, np1y, np2y, np1z, np2z, maxstp, nitera, b(6), f(4)
  double precision :: r(20), tfinal
  open(8, file=path)
  read(8, *) np1x, np2x
  read(8, *) np1y, np2y
  read(8, *) np1z, np2z
  read(8, *) r(1), r(2)
  read(8, *) r(3), r(4)
  read(8, *) r(5), r(6)
  read(8, *) r(7), r(8)
  read(8, *) r(9), r(10)
  read(8, *) r(11), r(12)
  read(8, *) r(13), r(14)
  read(8, *) r(15), r(16)
  read(8, *) r(17), r(18)
  read(8, *) r(19), r(20)
  read(8, *) maxstp, nitera, tfinal
  read(8, *) b
  read(8, *) f
  close(8)
  cfg%np1x = np1x; cfg%np2x = np2x; cfg%np1y = np1y; cfg%np2y = np2y; cfg%np1z = np1z; cfg%np2z = np2z
  cfg%ql1x = r(1); cfg%ql2x = r(2); cfg%ql1y = r(3); cfg%ql2y = r(4); cfg%ql1z = r(5); cfg%ql2z = r(6)
  cfg%vinitx1 = r(7); cfg%vinitx2 = r(8); cfg%vinity1 = r(9); cfg%vinity2 = r(10)
  cfg%vinitz1 = r(11); cfg%vinitz2 = r(12)
  cfg%binitx1 = r(13); cfg%binitx2 = r(14); cfg%binity1 = r(15); cfg%binity2 = r(16)
  cfg%binitz1 = r(17); cfg%binitz2 = r(18)
  cfg%cfl = r(19); cfg%facdep = r(20)
  cfg%maxstp = maxstp; cfg%nitera = nitera; cfg%tfinal = tfinal
  cfg%bcx1 = b(1); cfg%bcx2 = b(2); cfg%bcy1 = b(3); cfg%bcy2 = b(4); cfg%bcz1 = b(5); cfg%bcz2 = b(6)
  cfg%ifchem = f(1); cfg%ifthrm = f(2); cfg%ifrad = f(3); cfg%ifgrv = f(4)
end subroutine read_input3d


! INITIA's two-state x shock tube (M3:207-253)
subroutine shock_tube_ic(cfg, U, SP, nx, ny, nz)
  use mcf_abi_mod
  type(mcf_config), intent(in) :: cfg
  integer, intent(in) :: nx, ny, nz
  double precision, intent(out) :: U(-1:nx+2, -1:ny+2, -1:nz+2, 8), SP(-1:nx+2, -1:ny+2, -1:nz+2, 8)
  double precision :: dinit1, pinit1
  integer :: i
  dinit1 = cfg%Hini + cfg%pini + 2.d0*cfg%H2ini + 4.d0*cfg%Heini + 4.d0*cfg%Hepini
  pinit1 = cfg%pinit_over_kb*8.63359d0*1.d-3
  do i = -1, nx + 2
    U(i,:,:,1) = dinit1
    U(i,:,:,3) = cfg%vinity1
    U(i,:,:,4) = cfg%vinitz1
    U(i,:,:,5) = pinit1
    U(i,:,:,6) = cfg%binitx1
    U(i,:,:,8) = cfg%binitz1
    if (i <= cfg%np1x) then
      U(i,:,:,2) = cfg%vinitx1
      U(i,:,:,7) = cfg%binity1
    else
      U(i,:,:,2) = cfg%vinitx2
      U(i,:,:,7) = cfg%binity2
    end if
  end do
  SP(:,:,:,1) = cfg%Hini;  SP(:,:,:,2) = cfg%pini;  SP(:,:,:,3) = cfg%H2ini; SP(:,:,:,4) = cfg%Heini
  SP(:,:,:,5) = cfg%Hepini; SP(:,:,:,6) = cfg%Cini; SP(:,:,:,7) = cfg%COini; SP(:,:,:,8) = cfg%Cpini
end subroutine shock_tube_ic


! Restart files of an NPE-rank run into the global arrays (M3:454-465);
! ranks are read in order, so shared N+1 planes take the neighbour's value.
subroutine read_restart(dir, U, SP, nx, ny, nz, sx, sy, sz)
  character(*), intent(in) :: dir
  integer, intent(in) :: nx, ny, nz, sx, sy, sz
  double precision, intent(inout) :: U(-1:nx+2, -1:ny+2, -1:nz+2, 8), SP(-1:nx+2, -1:ny+2, -1:nz+2, 8)
  integer :: r, ist, jst, kst, cx, cy, cz, i, j, k, m
  double precision, allocatable :: rec(:,:)
  character(len=3) :: rs
  cx = nx/sx; cy = ny/sy; cz = nz/sz
  allocate(rec(17, cx + 1))
  do r = 0, sx*sy*sz - 1
    ist = mod(r, sx); kst = r/(sx*sy); jst = r/sx - sy*kst
    write(rs, '(I3.3)') r
    open(8, file=dir//'/000'//rs//'.dat', form='unformatted')
    do k = 1, cz + 1
      do j = 1, cy + 1
        read(8) rec
        do i = 1, cx + 1
          do m = 1, 8
            U(ist*cx + i, jst*cy + j, kst*cz + k, m) = rec(m, i)
            SP(ist*cx + i, jst*cy + j, kst*cz + k, m) = rec(8 + m, i)
          end do
        end do
      end do
    end do
    close(8)
  end do
end subroutine read_restart


! Per-rank FP32 snapshot {nunit}{rank}.dat and FP64 restart 000{rank}.dat
! (M3:701-717, 749-757); the 17th value is the potential Phi (zero without
! self-gravity).
subroutine write_rank_files(dir, U, SP, PH, nx, ny, nz, sx, sy, sz, nunit)
  character(*), intent(in) :: dir
  integer, intent(in) :: nx, ny, nz, sx, sy, sz, nunit
  double precision, intent(in) :: U(-1:nx+2, -1:ny+2, -1:nz+2, 8), SP(-1:nx+2, -1:ny+2, -1:nz+2, 8)
  double precision, intent(in) :: PH(-1:nx+2, -1:ny+2, -1:nz+2)
  integer :: r, ist, jst, kst, cx, cy, cz, i, j, k, m, gi, gj, gk
  double precision, allocatable :: rec(:,:)
  character(len=3) :: rs, ns
  cx = nx/sx; cy = ny/sy; cz = nz/sz
  allocate(rec(17, cx + 1))
  write(ns, '(I3.3)') nunit
  do r = 0, sx*sy*sz - 1
    ist = mod(r, sx); kst = r/(sx*sy); jst = r/sx - sy*kst
    write(rs, '(I3.3)') r
    open(10, file=dir//'/'//ns//rs//'.dat', form='unformatted')
    open(8, file=dir//'/000'//rs//'.dat', form='unformatted')
    do k = 1, cz + 1
      do j = 1, cy + 1
        gk = kst*cz + k; gj = jst*cy + j
        do i = 1, cx + 1
          gi = ist*cx + i
          do m = 1, 8
            rec(m, i) = U(gi, gj, gk, m)
            rec(8 + m, i) = SP(gi, gj, gk, m)
          end do
          rec(17, i) = PH(gi, gj, gk)
        end do
        write(10) real(rec, 4)
        write(8) rec
      end do
    end do
    close(10)
    close(8)
  end do
end subroutine write_rank_files


! cdnt.DAT (M3:300-307): the global cell centres along x, y and z, one
! D25.17 value per line, from the grid INITIA builds (M3:255-280)
subroutine write_cdnt(dir, cfg, nx, ny, nz)
  use mcf_abi_mod
  character(*), intent(in) :: dir
  type(mcf_config), intent(in) :: cfg
  integer, intent(in) :: nx, ny, nz
  double precision :: x_i(-1:nx+2), y_i(-1:ny+2), z_i(-1:nz+2), d(-1:max(nx, ny, nz)+2)
  integer :: i
  do i = -1, nx + 2
    if (i <= cfg%np1x) then
      d(i) = cfg%ql1x/dble(cfg%np1x)
    else
      d(i) = cfg%ql2x/dble(cfg%np2x)
    end if
  end do
  x_i(-1) = -d(0)
  do i = 0, nx + 2
    x_i(i) = x_i(i-1) + d(i)
  end do
  d = cfg%ql1y/dble(cfg%np1y)
  y_i(-1) = -d(0)
  do i = 0, ny + 2
    y_i(i) = y_i(i-1) + d(i)
  end do
  d = cfg%ql1z/dble(cfg%np1z)
  z_i(-1) = -d(0)
  do i = 0, nz + 2
    z_i(i) = z_i(i-1) + d(i)
  end do
  open(4, file=dir//'/cdnt.DAT')
  write(4, '(D25.17)') (0.5d0*(x_i(i-1) + x_i(i)), i = 1, nx)
  write(4, '(D25.17)') (0.5d0*(y_i(i-1) + y_i(i)), i = 1, ny)
  write(4, '(D25.17)') (0.5d0*(z_i(i-1) + z_i(i)), i = 1, nz)
  close(4)
end subroutine write_cdnt


! Per-rank 2D slices 2D{nunit2D}{rank}.dat (M3:763-783): the rank's local
! k = 1 plane, one FP32 record of 17 values per cell (rho, v, p, Bcc,
! species, Phi) for i = 1..Ncellx, per j = 1..Ncelly
subroutine write_rank_slices(dir, U, SP, BC, PH, nx, ny, nz, sx, sy, sz, nunit2D)
  character(*), intent(in) :: dir
  integer, intent(in) :: nx, ny, nz, sx, sy, sz, nunit2D
  double precision, intent(in) :: U(-1:nx+2, -1:ny+2, -1:nz+2, 8), SP(-1:nx+2, -1:ny+2, -1:nz+2, 8)
  double precision, intent(in) :: BC(-1:nx+2, -1:ny+2, -1:nz+2, 3), PH(-1:nx+2, -1:ny+2, -1:nz+2)
  integer :: r, ist, jst, kst, cx, cy, cz, i, j, m, gi, gj, gk
  real(4), allocatable :: rec(:,:)
  character(len=3) :: rs, ns
  cx = nx/sx; cy = ny/sy; cz = nz/sz
  allocate(rec(17, cx))
  write(ns, '(I3.3)') nunit2D
  do r = 0, sx*sy*sz - 1
    ist = mod(r, sx); kst = r/(sx*sy); jst = r/sx - sy*kst
    write(rs, '(I3.3)') r
    open(11, file=dir//'/2D'//ns//rs//'.dat', form='unformatted')
    gk = kst*cz + 1
    do j = 1, cy
      gj = jst*cy + j
      do i = 1, cx
        gi = ist*cx + i
        do m = 1, 5
          rec(m, i) = real(U(gi, gj, gk, m), 4)
        end do
        do m = 1, 3
          rec(5 + m, i) = real(BC(gi, gj, gk, m), 4)
        end do
        do m = 1, 8
          rec(8 + m, i) = real(SP(gi, gj, gk, m), 4)
        end do
        rec(17, i) = real(PH(gi, gj, gk), 4)
      end do
      write(11) rec
    end do
    close(11)
  end do
end subroutine write_rank_slices
