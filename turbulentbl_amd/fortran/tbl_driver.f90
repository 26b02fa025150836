!=======================================================================
! tbl_driver -- Fortran time loop of the TBL fractional step on the GPU.
!
! Mirrors PROGRAM revamptbl (dnssetbl/dnssetbl.f90:1-1269): parameters
! (here a namelist with the reference's names), mesh, initial condition,
! then the time loop calling each block of the step in the reference's
! order through the iso_c_binding module tbl_gpu, printing
! "time nstep |sum div|/N" every step like dnssetbl.f90:881.
!
! Usage: tbl_driver [namelist-file]     (namelist group /tbl/)
!=======================================================================
program tbl_driver
  use iso_c_binding
  use tbl_gpu
  implicit none
  integer :: nX = 90, nY = 80, nZ = 50, nsteps = 10, kmax = 50, mmax = 1000, gs_mmax = 1000
  integer :: freqTimAvg = 500, i_rec = 68, precision = 0, tol_mode = 0, precond = 4, helm_mode = 1
  integer :: blocks = 1, nprof = 37
  real(8) :: dt = 1.0d-4, Re = 150.92d0, Ue = 19.798d0, delta = 1.95d0, errorstop = 1.0d-6, errtol = 1.0d-5
  real(8) :: beta = 10.0d0, gama = 10.0d0, Lx = 12.0d0, Ly = 4.0d0, Lz = 3.4d0
  character(len=256) :: prof_file = 'Input/initProf.dat', out_file = ''
  namelist /tbl/ nX, nY, nZ, nsteps, dt, Re, Ue, delta, errorstop, gs_mmax, errtol, kmax, mmax, &
       freqTimAvg, i_rec, precision, tol_mode, precond, helm_mode, blocks, beta, gama, Lx, Ly, Lz, &
       prof_file, out_file, nprof

  type(tbl_params) :: p
  type(tbl_diag), allocatable :: dg(:)
  type(c_ptr) :: ctx
  real(8), allocatable, target :: Ut(:,:,:,:), Press(:,:,:), PCorr(:,:,:)
  real(8), allocatable, target :: x1(:), x2(:), x3(:), Jac(:), Jac2(:), Ab(:), At(:), Ap(:)
  real(8), allocatable, target :: avg(:,:,:), Wt(:,:), Ume(:), Ufluc(:,:,:)
  integer(c_int), allocatable, target :: newk(:)
  integer(c_int) :: sweeps(3), cycles, iters, guard
  real(8) :: time, div, mass(5), tres, ke, pert
  integer :: i, j, k, nstep, narg, u, ios
  character(len=256) :: nmlfile
  logical :: have

  narg = command_argument_count()
  if (narg >= 1) then
     call get_command_argument(1, nmlfile)
     open(newunit=u, file=trim(nmlfile), status='old', action='read')
     read(u, nml=tbl)
     close(u)
  end if

  call gpu_default_params(p)
  p%nX = nX; p%nY = nY; p%nZ = nZ; p%dt = dt; p%Re = Re; p%Ue = Ue; p%delta = delta
  p%Lx = Lx; p%Ly = Ly; p%Lz = Lz
  p%errorstop = errorstop; p%gs_mmax = gs_mmax; p%errtol = errtol; p%kmax = kmax; p%mmax = mmax
  p%freq_tim_avg = freqTimAvg; p%i_rec = i_rec; p%precision = precision; p%tol_mode = tol_mode
  p%precond = precond; p%helm_mode = helm_mode
  if (nX < 86) p%planes = [max(2, nint(5*nX/90.)), max(2, nint(20*nX/90.)), max(2, nint(40*nX/90.)), &
       max(2, nint(60*nX/90.)), min(nX-1, nint(85*nX/90.))]
  call gpu_create(p, ctx)

  allocate(Ut(nX,nY,nZ,3), Press(nX,nY,nZ), PCorr(nX,nY,nZ))
  allocate(x1(nX), x2(nY), x3(nZ), Jac(nZ), Jac2(nZ), Ab(nZ), At(nZ), Ap(nZ))
  allocate(avg(nX,nZ,9), Wt(nX,nZ), Ume(nZ), Ufluc(nY,nZ,3), newk(nZ))
  call gpu_mesh(ctx, x1, x2, x3, Jac, Jac2, Ab, At, Ap)

  ! inflow mean profile: Input/initProf.dat (24x,E18.8) as dnssetbl.f90:106-115,
  ! else a Reichardt law in wall units capped at Ue
  inquire(file=trim(prof_file), exist=have)
  if (have) then
     open(newunit=u, file=trim(prof_file), status='old', action='read')
     do k = 1, min(nprof, nZ)
        read(u, '(24x,E18.8)', iostat=ios) Ume(k)
     end do
     close(u)
     do k = nprof + 1, nZ
        Ume(k) = Ume(nprof)
     end do
  else
     do k = 1, nZ
        Ume(k) = min(Ue, log(1d0 + 0.41d0*max(x3(k),0d0)*Re)/0.41d0 + &
             7.8d0*(1d0 - exp(-max(x3(k),0d0)*Re/11d0) - (max(x3(k),0d0)*Re/11d0)*exp(-max(x3(k),0d0)*Re/3d0)))
     end do
  end if

  ! initial field of dnssetbl.f90:151-166 (prmpt = 2), with the mesh already set
  do k = 1, nZ
     do j = 1, nY
        do i = 1, nX
           pert = 0.01d0*Ue*sin(beta*x2(j))*cos(gama*x3(k))
           Ut(i,j,k,1) = Ume(k) + pert
           Ut(i,j,k,2) = pert
           Ut(i,j,k,3) = pert
        end do
     end do
  end do
  Press = 0d0
  PCorr = 0d0
  call gpu_upload(ctx, Ut, Press, PCorr)
  avg = 0d0; Wt = 0d0; Ufluc = 0d0
  do k = 1, nZ
     newk(k) = min(k, nZ - 1)
  end do
  call gpu_set_recycle(ctx, avg, Wt, newk, Ume, Ufluc, 1d0, 1d0, 0d0, 0)

  time = 0d0
  allocate(dg(1))
  do nstep = 1, nsteps
     time = time + dt
     if (blocks /= 0) then
        call begin_step(ctx)
        call recycle_inflow(ctx)
        call momentum_rhs(ctx)
        call diffusion_solve(ctx, errorstop, gs_mmax, sweeps)
        call ppe_rhs(ctx)
        call ppe_solve(ctx, errtol, kmax, mmax, cycles, iters, tres)
        call project(ctx)
        call apply_bcs(ctx)
        call div_check(ctx, div)
        call mass_flux(ctx, mass)
        call stats_accumulate(ctx)
        if (mod(nstep, freqTimAvg) == 0) call rescale(ctx)
        if (nstep > freqTimAvg) then
           call ke_guard(ctx, ke, guard)
           if (guard /= 0) then
              write(*,*) 'Very Large Fluctuations'
              stop 2
           end if
        end if
     else
        call time_step(ctx, 1, dg)
        div = dg(1)%div
     end if
     write(*,105) time, nstep, div
  end do
105 format(1X,F14.7,1X,I9,1X,F14.7)

  call gpu_download(ctx, Ut, Press, PCorr)
  if (len_trim(out_file) > 0) then
     open(newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
     write(u) Ut, Press
     close(u)
  end if
  write(*,'(a,es24.16)') ' checksum ', sum(Ut(2:nX-1,2:nY-1,2:nZ-1,:))
  call gpu_destroy(ctx)
end program tbl_driver
