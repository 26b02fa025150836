!=======================================================================
! tbl_driver -- Fortran time loop of the TBL fractional step on the GPU.
!
! Mirrors PROGRAM revamptbl (dnssetbl/dnssetbl.f90:1-1269): parameters
! (here a namelist with the reference's names), mesh, initial condition,
! then the time loop calling each block of the step in the reference's
! order through the iso_c_binding module tbl_gpu, printing
! "time nstep |sum div|/N" every step like dnssetbl.f90:881.
!
! Restart (prmpt = 1) reads Input/3D.dat and Input/statData.dat in the
! reference's formats (dnssetbl.f90:118-150); prmpt = 2 starts from the
! initial condition of :151-166.  dump_file writes the final field in the
! combined restart format (combineFile.f90:55-71), which prmpt = 1 reads back.
! out_dir: at every statistics rescale the driver writes the reference's
! Output files lamda.dat (appended), W.dat, boundLyrThik.dat, newK.dat,
! RMS.dat and meanVel.dat there, and statData.dat in the format of
! Input/statData.dat (dnssetbl.f90:1071-1198).
! tau_window > 0: the wall shear of every step is summed and, at each step
! with mod(nstep, tau_window) == 0, the window mean per station is appended to
! out_dir/numrcltauwall.xy as time, x, tau/(nY-2), C_f in format 94
! (dnstbl/dnstbl.f90:1666-1682, 2389).
!
! ckpt_in / ckpt_out: resume from / write a binary checkpoint
! (tbl_checkpoint_load / _save; bitwise resume).  With ckpt_in the fields,
! statistics, recycling state and clock come from the checkpoint.
!
! Usage: tbl_driver [namelist-file]     (namelist group /tbl/)
!=======================================================================
program tbl_driver
  use iso_c_binding
  use tbl_gpu
  implicit none
  integer :: nX = 90, nY = 80, nZ = 50, nsteps = 10, kmax = 50, mmax = 1000, gs_mmax = 1000
  integer :: freqTimAvg = 500, i_rec = 68, precision = 0, tol_mode = 0, precond = 4, helm_mode = 3
  integer :: blocks = 1, nprof = 37, prmpt = 2, scheme = 0, tau_window = 0, slabs = 1
  real(8) :: forc(3) = 0d0
  real(8) :: dt = 1.0d-4, Re = 150.92d0, Ue = 19.798d0, delta = 1.95d0, errorstop = 1.0d-6, errtol = 1.0d-5
  real(8) :: beta = 10.0d0, gama = 10.0d0, Lx = 12.0d0, Ly = 4.0d0, Lz = 3.4d0
  character(len=256) :: prof_file = 'Input/initProf.dat', out_file = ''
  character(len=256) :: in_file = 'Input/3D.dat', stat_file = 'Input/statData.dat', dump_file = ''
  character(len=256) :: out_dir = '', ckpt_in = '', ckpt_out = ''
  namelist /tbl/ nX, nY, nZ, nsteps, dt, Re, Ue, delta, errorstop, gs_mmax, errtol, kmax, mmax, &
       freqTimAvg, i_rec, precision, tol_mode, precond, helm_mode, blocks, beta, gama, Lx, Ly, Lz, &
       prof_file, out_file, nprof, prmpt, in_file, stat_file, dump_file, out_dir, scheme, forc, tau_window, &
       ckpt_in, ckpt_out, slabs

  type(tbl_params) :: p
  type(tbl_diag), allocatable :: dg(:)
  type(c_ptr) :: ctx
  real(8), allocatable, target :: Ut(:,:,:,:), Press(:,:,:), PCorr(:,:,:)
  real(8), allocatable, target :: x1(:), x2(:), x3(:), Jac(:), Jac2(:), Ab(:), At(:), Ap(:)
  real(8), allocatable, target :: avg(:,:,:), Wt(:,:), Ume(:), Ufluc(:,:,:)
  real(8), allocatable, target :: tau(:), sumxz(:), taum(:), cfx(:)
  integer(c_int) :: nwin
  integer(c_int), allocatable, target :: newk(:)
  integer(c_int) :: sweeps(3), cycles, iters, guard
  real(8) :: time, div, mass(5), tres, ke, pert
  integer :: i, j, k, nstep, narg, u, ios, nstep0, cnt0
  real(8) :: xr1, xr2, xr3, usum2k, vsum2k, wsum2k, wt1k
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
  p%scheme = scheme; p%forc = forc; p%slabs = slabs
  p%wall_shear_window = tau_window
  if (nX < 86) p%planes = [max(2, nint(5*nX/90.)), max(2, nint(20*nX/90.)), max(2, nint(40*nX/90.)), &
       max(2, nint(60*nX/90.)), min(nX-1, nint(85*nX/90.))]
  call gpu_create(p, ctx)

  allocate(Ut(nX,nY,nZ,3), Press(nX,nY,nZ), PCorr(nX,nY,nZ))
  allocate(x1(nX), x2(nY), x3(nZ), Jac(nZ), Jac2(nZ), Ab(nZ), At(nZ), Ap(nZ))
  allocate(avg(nX,nZ,9), Wt(nX,nZ), Ume(nZ), Ufluc(nY,nZ,3), newk(nZ))
  allocate(tau(nX), sumxz(nX), taum(nX), cfx(nX))
  sumxz = 0d0
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

  avg = 0d0; Wt = 0d0; Ufluc = 0d0
  do k = 1, nZ
     newk(k) = min(k, nZ - 1)
  end do
  Press = 0d0
  PCorr = 0d0
  time = 0d0
  nstep0 = 0
  if (prmpt == 1) then
     ! restart: the combined field file and the recycling-plane statistics
     open(newunit=u, file=trim(in_file), status='old', action='read')
     read(u, 101) time, nstep0
     read(u, *)
     read(u, *)
     do i = 1, nX
        read(u, *)
        do j = 1, nY
           read(u, *)
           do k = 1, nZ
              read(u, 104) xr1, xr2, xr3, Ut(i,j,k,1), Ut(i,j,k,2), Ut(i,j,k,3), Press(i,j,k)
           end do
        end do
     end do
     close(u)
     open(newunit=u, file=trim(stat_file), status='old', action='read')
     do k = 1, nZ
        read(u, 115) usum2k, vsum2k, wsum2k, wt1k, newk(k)
        avg(i_rec, k, 1) = usum2k
        avg(i_rec, k, 2) = vsum2k
        avg(i_rec, k, 3) = wsum2k
        Wt(1, k) = wt1k
     end do
     close(u)
     do k = 2, nZ - 1
        do j = 2, nY - 1
           Ufluc(j, k, 1) = Ut(i_rec, j, k, 1) - avg(i_rec, k, 1)
           Ufluc(j, k, 2) = Ut(i_rec, j, k, 2) - avg(i_rec, k, 2)
           Ufluc(j, k, 3) = Ut(i_rec, j, k, 3) - avg(i_rec, k, 3)
        end do
     end do
  else
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
  end if
  call gpu_upload(ctx, Ut, Press, PCorr)
  call gpu_set_recycle(ctx, avg, Wt, newk, Ume, Ufluc, 1d0, 1d0, time, nstep0)
  cnt0 = 0
  if (len_trim(ckpt_in) > 0) then
     call checkpoint_load(ctx, ckpt_in)
     call gpu_clock(ctx, time, nstep0, cnt0)
  end if
101 FORMAT(8X,F20.10,I9,1X)
104 FORMAT(3(1X,F10.6),4(1X,E22.15))
115 FORMAT(4(1X,F15.7),I5.3)

  allocate(dg(1))
  do nstep = nstep0 + 1, nstep0 + nsteps
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
        if (mod(nstep, freqTimAvg) == 0) then
           call rescale(ctx)
           if (len_trim(out_dir) > 0) call write_rescale_files()
        end if
        if (tau_window > 0) then
           ! the reference's host-side window sum (dnstbl/dnstbl.f90:1666-1682)
           call wall_shear(ctx, tau)
           sumxz(2:nX-1) = tau(2:nX-1) + sumxz(2:nX-1)
           if (mod(nstep, tau_window) == 0) then
              sumxz = sumxz / dble(tau_window)
              taum = sumxz / dble(nY - 2)
              cfx = 2.0d0 * (sumxz / (Ue * Ue * dble(nY - 2)))
              call write_tau_file()
              sumxz = 0d0
           end if
        end if
        if (cnt0 + nstep - nstep0 > freqTimAvg) then   ! cnt of dnssetbl.f90:1223 (steps since the run began)
           call ke_guard(ctx, ke, guard)
           if (guard /= 0) then
              write(*,*) 'Very Large Fluctuations'
              stop 2
           end if
        end if
     else
        call time_step(ctx, 1, dg)
        div = dg(1)%div
        if (dg(1)%cf_ready /= 0) then
           call wall_shear_mean(ctx, taum, cfx, nwin)
           call write_tau_file()
        end if
     end if
     write(*,105) time, nstep, div
  end do
105 format(1X,F14.7,1X,I9,1X,F14.7)

  if (len_trim(ckpt_out) > 0) call checkpoint_save(ctx, ckpt_out)
  call gpu_download(ctx, Ut, Press, PCorr)
  if (len_trim(out_file) > 0) then
     open(newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
     write(u) Ut, Press
     close(u)
  end if
  if (len_trim(dump_file) > 0) then
     ! combined restart format (dnssetbl/combineFile.f90:55-71): all span rows
     open(newunit=u, file=trim(dump_file), status='replace', action='write')
     write(u, 205) 'TITLE ="', time, nstep0 + nsteps, '"'
     write(u, '(a)') 'Variables = "x","y","z","U","V","W","P"'
     write(u, 206) 'ZONE k=', nX, ',j=', nY, ',i=', nZ, ',DATAPACKING=POINT'
     do i = 1, nX
        write(u, *)
        do j = 1, nY
           write(u, *)
           do k = 1, nZ
              write(u, 104) x1(i), x2(j), x3(k), Ut(i,j,k,1), Ut(i,j,k,2), Ut(i,j,k,3), Press(i,j,k)
           end do
        end do
     end do
     close(u)
  end if
205 format(A,F20.10,I9,A)
206 format(A,I4,A,I4,A,I4,A)
  write(*,'(a,es24.16)') ' checksum ', sum(Ut(2:nX-1,2:nY-1,2:nZ-1,:))
  call gpu_destroy(ctx)

contains

  ! the files of dnssetbl.f90:1071-1198 from the statistics the library holds
  ! after tbl_rescale (Re_x is never assigned by the reference: written as 0;
  ! eta(nZ) likewise)
  subroutine write_rescale_files()
    real(8), allocatable, target :: sav(:,:,:), sWt(:,:), bl(:), ds(:), th(:)
    integer(c_int), allocatable, target :: nk(:)
    real(8) :: lam, lam03, eta(nZ), rr(6)
    integer :: ii, kk, v
    allocate(sav(nX,nZ,9), sWt(nX,nZ), bl(nX), ds(nX), th(nX), nk(nZ))
    call gpu_get_stats(ctx, sav, sWt, nk, bl, ds, th, lam, lam03)
    open(newunit=v, file=trim(out_dir)//'/lamda.dat', position='append', action='write')
    write(v, 110) time, lam, lam03
    close(v)
    open(newunit=v, file=trim(out_dir)//'/W.dat', status='replace', action='write')
    write(v, *) 'Variables = "x1","x3","Wt"'
    do ii = 1, nX
       do kk = 1, nZ
          write(v, 111) x1(ii), x3(kk), sWt(ii,kk)
       end do
    end do
    close(v)
    open(newunit=v, file=trim(out_dir)//'/boundLyrThik.dat', status='replace', action='write')
    write(v, *) 'Variables = "x1","BLthick","deltastar","theta","Ht","Reth","Rex"'
    do ii = 1, nX
       write(v, 112) x1(ii), bl(ii), ds(ii), th(ii), ds(ii)/th(ii), th(ii)*Re*Ue, 0d0
    end do
    close(v)
    eta = 0d0
    do kk = 1, nZ - 1
       if (sWt(i_rec,kk) <= 0.99d0) then
          eta(kk) = lam
       else
          eta(kk) = lam03
       end if
    end do
    open(newunit=v, file=trim(out_dir)//'/newK.dat', status='replace', action='write')
    write(v, *) 'Variables = "k","newk","etak"'
    do kk = 1, nZ
       write(v, 113) kk, nk(kk), eta(kk)
    end do
    close(v)
    open(newunit=v, file=trim(out_dir)//'/RMS.dat', status='replace', action='write')
    write(v, *) 'Variables = "x1","x3","Ruu","Rvv","Rww","Ruv","Ruw","Rvw"'
    do ii = 1, nX
       do kk = 1, nZ
          rr(1) = sav(ii,kk,4) - sav(ii,kk,1)*sav(ii,kk,1)
          rr(2) = sav(ii,kk,5) - sav(ii,kk,2)*sav(ii,kk,2)
          rr(3) = sav(ii,kk,6) - sav(ii,kk,3)*sav(ii,kk,3)
          rr(4) = sav(ii,kk,7) - sav(ii,kk,1)*sav(ii,kk,2)
          rr(5) = sav(ii,kk,8) - sav(ii,kk,1)*sav(ii,kk,3)
          rr(6) = sav(ii,kk,9) - sav(ii,kk,2)*sav(ii,kk,3)
          write(v, 116) x1(ii), x3(kk), rr
       end do
    end do
    close(v)
    open(newunit=v, file=trim(out_dir)//'/meanVel.dat', status='replace', action='write')
    write(v, *) 'Variables = "x1","x3","usum2","vsum2","wsum2"'
    do ii = 1, nX
       do kk = 1, nZ
          write(v, 114) x1(ii), x3(kk), sav(ii,kk,1), sav(ii,kk,2), sav(ii,kk,3)
       end do
    end do
    close(v)
    open(newunit=v, file=trim(out_dir)//'/statData.dat', status='replace', action='write')
    do kk = 1, nZ
       write(v, 115) sav(i_rec,kk,1), sav(i_rec,kk,2), sav(i_rec,kk,3), sWt(1,kk), nk(kk)
    end do
    close(v)
110 format(3(2x,F20.10))
111 format(3(1X,F15.7))
112 format(8(2x,F20.10))
113 format(2(I5),F20.10)
114 format(5(1X,F15.7))
116 format(8(1X,F15.7))
115 format(4(1X,F15.7),I5.3)
  end subroutine write_rescale_files

  ! the window mean wall shear per station, appended as the reference does
  ! (dnstbl/dnstbl.f90:1674-1678, format 94 at :2389)
  subroutine write_tau_file()
    integer :: uw, ii
    if (len_trim(out_dir) == 0) return
    open(newunit=uw, file=trim(out_dir)//'/numrcltauwall.xy', status='unknown', position='append')
    do ii = 2, nX - 1
       write(uw, 94) time, x1(ii), taum(ii), cfx(ii)
    end do
    close(uw)
94  format(4(1X,F15.7))
  end subroutine write_tau_file
end program tbl_driver
