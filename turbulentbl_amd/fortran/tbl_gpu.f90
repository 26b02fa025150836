!=======================================================================
! tbl_gpu -- iso_c_binding layer over libtbl_hip.so (include/tbl.h).
!
! The reference (dnssetbl/dnssetbl.f90) runs every block of its time step as
! an inline loop nest over global arrays.  This module gives each block a
! subroutine with the reference's array shapes -- Ut(nX,nY,nZ,3),
! Press(nX,nY,nZ), PCorr(nX,nY,nZ), statistics (nX,nZ) -- that forwards to
! the HIP library, so a Fortran time loop can call them in the reference's
! order (dnssetbl.f90:253-1262) with the fields resident on the GPU.
!=======================================================================
module tbl_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: TBL_FP64 = 0, TBL_MIXED = 1
  integer(c_int), parameter, public :: TBL_PC_SSOR_RB = 1, TBL_PC_NONE = 2, TBL_PC_JACOBI = 3, TBL_PC_FAST = 4, &
       TBL_PC_SSOR_BLOCK = 5
  integer(c_int), parameter, public :: TBL_TOL_REF = 0, TBL_TOL_TRUE = 1
  integer(c_int), parameter, public :: TBL_HELM_POINT_RB = 0, TBL_HELM_ZLINE_RB = 1, TBL_HELM_DIRECT = 2, &
       TBL_HELM_ZEBRA = 3, TBL_HELM_LEX_SLAB = 4
  integer(c_int), parameter, public :: TBL_SCHEME_SETBL = 0, TBL_SCHEME_DNSTBL = 1

  ! layout of struct tbl_params
  type, bind(C), public :: tbl_params
     integer(c_int) :: nX, nY, nZ
     real(c_double) :: Lx, Ly, Lz, delta, dt, Re, Ue, omega
     integer(c_int) :: i_rec, freq_tim_avg
     integer(c_int) :: planes(5)
     integer(c_int) :: precision
     real(c_double) :: errorstop
     integer(c_int) :: gs_mmax, helm_mode, tol_mode
     real(c_double) :: errtol
     integer(c_int) :: kmax, mmax, precond
     real(c_double) :: inner_tol
     integer(c_int) :: fail_on_noconv, wall_shear_window, pc_extrap
     integer(c_int) :: scheme
     real(c_double) :: forc(3)
     integer(c_int) :: slabs
  end type tbl_params

  ! layout of struct tbl_state
  type, bind(C), public :: tbl_state
     type(c_ptr) :: sums = c_null_ptr, avg = c_null_ptr, Wt = c_null_ptr, newk = c_null_ptr, Ume = c_null_ptr
     type(c_ptr) :: Ufluc_rec = c_null_ptr, bltkns = c_null_ptr, deltastar = c_null_ptr, theta = c_null_ptr
     type(c_ptr) :: wall_w = c_null_ptr
     real(c_double) :: lamda = 1d0, lamda03 = 1d0, time = 0d0
     integer(c_int) :: nstep = 0, cnt = 0
  end type tbl_state

  ! layout of struct tbl_adjoint_opts
  type, bind(C), public :: tbl_adjoint_opts
     integer(c_int) :: nsteps, nsnap, i0, i1, forward_only, branch_mode
  end type tbl_adjoint_opts

  ! layout of struct tbl_diag
  type, bind(C), public :: tbl_diag
     real(c_double) :: div, div_signed, mass(5), ke
     integer(c_int) :: guard, gs_sweeps(3), gm_cycles, gm_iters
     real(c_double) :: gm_true_res, gm_sumav
     integer(c_int) :: rescaled, nstep, cf_ready
  end type tbl_diag

  interface
     subroutine c_default_params(p) bind(C, name='tbl_default_params')
       import :: tbl_params
       type(tbl_params), intent(out) :: p
     end subroutine
     integer(c_int) function c_create(p, ctx) bind(C, name='tbl_create')
       import :: tbl_params, c_int, c_ptr
       type(tbl_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function c_destroy(ctx) bind(C, name='tbl_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_last_error(ctx, buf, n) bind(C, name='tbl_last_error')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: n
     end function
     integer(c_int) function c_upload(ctx, Ut, P, PC) bind(C, name='tbl_upload')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, Ut, P, PC
     end function
     integer(c_int) function c_download(ctx, Ut, P, PC) bind(C, name='tbl_download')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, Ut, P, PC
     end function
     integer(c_int) function c_set_state(ctx, s) bind(C, name='tbl_set_state')
       import :: c_int, c_ptr, tbl_state
       type(c_ptr), value :: ctx
       type(tbl_state), intent(in) :: s
     end function
     integer(c_int) function c_get_state(ctx, s) bind(C, name='tbl_get_state')
       import :: c_int, c_ptr, tbl_state
       type(c_ptr), value :: ctx
       type(tbl_state), intent(inout) :: s
     end function
     integer(c_int) function c_get_mesh(ctx, x1, x2, x3, Jac, Jac2, Ab, At, Ap) bind(C, name='tbl_get_mesh')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, x1, x2, x3, Jac, Jac2, Ab, At, Ap
     end function
     integer(c_int) function c_begin_step(ctx) bind(C, name='tbl_begin_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_recycle_inflow(ctx) bind(C, name='tbl_recycle_inflow')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_momentum_rhs(ctx) bind(C, name='tbl_momentum_rhs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_diffusion_solve(ctx, tol, mmax, sweeps, res) bind(C, name='tbl_diffusion_solve')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: tol
       integer(c_int), value :: mmax
       integer(c_int) :: sweeps(3)
       real(c_double) :: res(3)
     end function
     integer(c_int) function c_ppe_rhs(ctx) bind(C, name='tbl_ppe_rhs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_ppe_solve(ctx, errtol, kmax, mmax, cycles, iters, tres) bind(C, name='tbl_ppe_solve')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: errtol
       integer(c_int), value :: kmax, mmax
       integer(c_int) :: cycles, iters
       real(c_double) :: tres
     end function
     integer(c_int) function c_project(ctx) bind(C, name='tbl_project')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_apply_bcs(ctx) bind(C, name='tbl_apply_bcs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_div_check(ctx, div, sdiv) bind(C, name='tbl_div_check')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: div, sdiv
     end function
     integer(c_int) function c_mass_flux(ctx, mass) bind(C, name='tbl_mass_flux')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mass(5)
     end function
     integer(c_int) function c_stats_accumulate(ctx) bind(C, name='tbl_stats_accumulate')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_rescale(ctx) bind(C, name='tbl_rescale')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function c_ke_guard(ctx, ke, guard) bind(C, name='tbl_ke_guard')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: ke
       integer(c_int) :: guard
     end function
     integer(c_int) function c_wall_shear(ctx, tau) bind(C, name='tbl_wall_shear')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: tau(*)
     end function
     integer(c_int) function c_wall_shear_mean(ctx, tau_mean, cf, nwin) bind(C, name='tbl_wall_shear_mean')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: tau_mean(*), cf(*)
       integer(c_int) :: nwin
     end function
     integer(c_int) function c_checkpoint_save(ctx, path) bind(C, name='tbl_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char) :: path(*)
     end function
     integer(c_int) function c_checkpoint_load(ctx, path) bind(C, name='tbl_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char) :: path(*)
     end function
     integer(c_int) function c_step(ctx, n, diag) bind(C, name='tbl_step')
       import :: c_int, c_ptr, tbl_diag
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       type(tbl_diag) :: diag(*)
     end function
     integer(c_int) function c_adjoint_window(ctx, o, wall_w, J, dJdw, nfwd, nadj) bind(C, name='tbl_adjoint_window')
       import :: c_int, c_ptr, c_double, tbl_adjoint_opts
       type(c_ptr), value :: ctx
       type(tbl_adjoint_opts), intent(in) :: o
       real(c_double), intent(in) :: wall_w(*)
       real(c_double), intent(out) :: J
       real(c_double) :: dJdw(*)
       integer(c_int), intent(out) :: nfwd, nadj
     end function
  end interface

  public :: gpu_default_params, gpu_create, gpu_destroy, gpu_check
  public :: gpu_upload, gpu_download, gpu_set_recycle, gpu_get_stats, gpu_mesh, gpu_clock
  public :: begin_step, recycle_inflow, momentum_rhs, diffusion_solve, ppe_rhs, ppe_solve
  public :: project, apply_bcs, div_check, mass_flux, stats_accumulate, rescale, ke_guard
  public :: wall_shear, wall_shear_mean, time_step, adjoint_window, checkpoint_save, checkpoint_load

contains

  subroutine gpu_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(256)
    integer :: i, ierr
    character(len=256) :: msg
    if (rc == 0) return
    ! a null ctx (a failed tbl_create) gives that call's message
    msg = ''
    ierr = c_last_error(ctx, buf, 256)
    do i = 1, 256
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*,'(a,a,a,i0,a,a)') 'tbl_gpu: ', where, ' failed, rc=', rc, ': ', trim(msg)
    stop 1
  end subroutine gpu_check

  subroutine gpu_default_params(p)
    type(tbl_params), intent(out) :: p
    call c_default_params(p)
  end subroutine gpu_default_params

  subroutine gpu_create(p, ctx)
    type(tbl_params), intent(in) :: p
    type(c_ptr), intent(out) :: ctx
    call gpu_check(c_null_ptr, c_create(p, ctx), 'tbl_create')
  end subroutine gpu_create

  subroutine gpu_destroy(ctx)
    type(c_ptr), intent(inout) :: ctx
    integer(c_int) :: rc
    rc = c_destroy(ctx)
    ctx = c_null_ptr
  end subroutine gpu_destroy

  ! restart data in (dnssetbl.f90:118-135) and solution out (:908-933)
  subroutine gpu_upload(ctx, Ut, Press, PCorr)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), target, contiguous :: Ut(:,:,:,:), Press(:,:,:), PCorr(:,:,:)
    call gpu_check(ctx, c_upload(ctx, c_loc(Ut), c_loc(Press), c_loc(PCorr)), 'tbl_upload')
  end subroutine gpu_upload

  subroutine gpu_download(ctx, Ut, Press, PCorr)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), target, contiguous :: Ut(:,:,:,:), Press(:,:,:), PCorr(:,:,:)
    call gpu_check(ctx, c_download(ctx, c_loc(Ut), c_loc(Press), c_loc(PCorr)), 'tbl_download')
  end subroutine gpu_download

  ! recycling state read from Input/statData.dat in the reference (:137-150)
  subroutine gpu_set_recycle(ctx, avg, Wt, newk, Ume, Ufluc_rec, lamda, lamda03, time, nstep)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in), target, contiguous :: avg(:,:,:), Wt(:,:), Ume(:), Ufluc_rec(:,:,:)
    integer(c_int), intent(in), target, contiguous :: newk(:)
    real(c_double), intent(in) :: lamda, lamda03, time
    integer, intent(in) :: nstep
    type(tbl_state) :: s
    s%avg = c_loc(avg)
    s%Wt = c_loc(Wt)
    s%newk = c_loc(newk)
    s%Ume = c_loc(Ume)
    s%Ufluc_rec = c_loc(Ufluc_rec)
    s%lamda = lamda
    s%lamda03 = lamda03
    s%time = time
    s%nstep = nstep
    s%cnt = 0
    call gpu_check(ctx, c_set_state(ctx, s), 'tbl_set_state')
  end subroutine gpu_set_recycle

  ! statistics after a rescale (dnssetbl.f90:1017-1145): the nine averages
  ! usum2..vwsum2 (nX,nZ,9), the Lund weight, new(k), delta99, delta*, theta,
  ! lamda and lamda03
  subroutine gpu_get_stats(ctx, avg, Wt, newk, bltkns, deltastar, theta, lamda, lamda03)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), target, contiguous :: avg(:,:,:), Wt(:,:), bltkns(:), deltastar(:), theta(:)
    integer(c_int), intent(out), target, contiguous :: newk(:)
    real(c_double), intent(out) :: lamda, lamda03
    type(tbl_state) :: s
    s%avg = c_loc(avg)
    s%Wt = c_loc(Wt)
    s%newk = c_loc(newk)
    s%bltkns = c_loc(bltkns)
    s%deltastar = c_loc(deltastar)
    s%theta = c_loc(theta)
    call gpu_check(ctx, c_get_state(ctx, s), 'tbl_get_state')
    lamda = s%lamda
    lamda03 = s%lamda03
  end subroutine gpu_get_stats

  ! the context's clock: time, step counter and the reference's cnt (steps of
  ! this run, dnssetbl.f90:256,1223)
  subroutine gpu_clock(ctx, time, nstep, cnt)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: time
    integer, intent(out) :: nstep, cnt
    type(tbl_state) :: s
    call gpu_check(ctx, c_get_state(ctx, s), 'tbl_get_state')
    time = s%time
    nstep = s%nstep
    cnt = s%cnt
  end subroutine gpu_clock

  ! mesh and coefficients (dnssetbl.f90:183-247), computed by the library
  subroutine gpu_mesh(ctx, x1, x2, x3, Jac, Jac2, Ab, At, Ap)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), target, contiguous :: x1(:), x2(:), x3(:), Jac(:), Jac2(:), Ab(:), At(:), Ap(:)
    call gpu_check(ctx, c_get_mesh(ctx, c_loc(x1), c_loc(x2), c_loc(x3), c_loc(Jac), c_loc(Jac2), &
         c_loc(Ab), c_loc(At), c_loc(Ap)), 'tbl_get_mesh')
  end subroutine gpu_mesh

  ! ---- the blocks of one step, in the reference's order ----
  subroutine begin_step(ctx)                 ! :254-257
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_begin_step(ctx), 'tbl_begin_step')
  end subroutine begin_step
  subroutine recycle_inflow(ctx)             ! :259-303
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_recycle_inflow(ctx), 'tbl_recycle_inflow')
  end subroutine recycle_inflow
  subroutine momentum_rhs(ctx)               ! :306-369
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_momentum_rhs(ctx), 'tbl_momentum_rhs')
  end subroutine momentum_rhs
  subroutine diffusion_solve(ctx, errorstop, mmax, sweeps)   ! :371-433
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in) :: errorstop
    integer, intent(in) :: mmax
    integer(c_int), intent(out) :: sweeps(3)
    real(c_double) :: res(3)
    call gpu_check(ctx, c_diffusion_solve(ctx, errorstop, int(mmax, c_int), sweeps, res), 'tbl_diffusion_solve')
  end subroutine diffusion_solve
  subroutine ppe_rhs(ctx)                    ! :437-447
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_ppe_rhs(ctx), 'tbl_ppe_rhs')
  end subroutine ppe_rhs
  subroutine ppe_solve(ctx, errtol, kmax, mmax, cycles, iters, true_res)   ! :449-671
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in) :: errtol
    integer, intent(in) :: kmax, mmax
    integer(c_int), intent(out) :: cycles, iters
    real(c_double), intent(out) :: true_res
    call gpu_check(ctx, c_ppe_solve(ctx, errtol, int(kmax, c_int), int(mmax, c_int), cycles, iters, true_res), &
         'tbl_ppe_solve')
  end subroutine ppe_solve
  subroutine project(ctx)                    ! :673-726
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_project(ctx), 'tbl_project')
  end subroutine project
  subroutine apply_bcs(ctx)                  ! :728-797
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_apply_bcs(ctx), 'tbl_apply_bcs')
  end subroutine apply_bcs
  subroutine div_check(ctx, div)             ! :851-882
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: div
    real(c_double) :: sdiv
    call gpu_check(ctx, c_div_check(ctx, div, sdiv), 'tbl_div_check')
  end subroutine div_check
  subroutine mass_flux(ctx, mass)            ! :885-904
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: mass(5)
    call gpu_check(ctx, c_mass_flux(ctx, mass), 'tbl_mass_flux')
  end subroutine mass_flux
  subroutine stats_accumulate(ctx)           ! :936-984
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_stats_accumulate(ctx), 'tbl_stats_accumulate')
  end subroutine stats_accumulate
  subroutine rescale(ctx)                    ! :986-1220
    type(c_ptr), intent(in) :: ctx
    call gpu_check(ctx, c_rescale(ctx), 'tbl_rescale')
  end subroutine rescale
  subroutine ke_guard(ctx, ke, guard)        ! :1223-1261
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: ke
    integer(c_int), intent(out) :: guard
    call gpu_check(ctx, c_ke_guard(ctx, ke, guard), 'tbl_ke_guard')
  end subroutine ke_guard
  subroutine wall_shear(ctx, tau)            ! dnstbl/dnstbl.f90:1620-1668
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: tau(:)
    call gpu_check(ctx, c_wall_shear(ctx, tau), 'tbl_wall_shear')
  end subroutine wall_shear
  subroutine wall_shear_mean(ctx, tau_mean, cf, nwin)   ! dnstbl/dnstbl.f90:1666-1682
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: tau_mean(:), cf(:)
    integer(c_int), intent(out) :: nwin
    call gpu_check(ctx, c_wall_shear_mean(ctx, tau_mean, cf, nwin), 'tbl_wall_shear_mean')
  end subroutine wall_shear_mean
  ! binary checkpoint of the device state; bitwise resume (replaces the
  ! E22.15 restart files of dnssetbl.f90:118-150 for resuming a run)
  subroutine checkpoint_save(ctx, path)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: path
    call gpu_check(ctx, c_checkpoint_save(ctx, trim(path)//c_null_char), 'tbl_checkpoint_save')
  end subroutine checkpoint_save
  subroutine checkpoint_load(ctx, path)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: path
    call gpu_check(ctx, c_checkpoint_load(ctx, trim(path)//c_null_char), 'tbl_checkpoint_load')
  end subroutine checkpoint_load
  subroutine time_step(ctx, n, diag)         ! whole steps, :253-1262
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n
    type(tbl_diag), intent(out) :: diag(n)
    call gpu_check(ctx, c_step(ctx, int(n, c_int), diag), 'tbl_step')
  end subroutine time_step

  ! wall-shear objective of a window of steps and its gradient with respect to
  ! the wall-normal wall velocity w(i,j,1) (the hook of dnssetbl.f90:729)
  subroutine adjoint_window(ctx, nsteps, nsnap, i0, i1, wall_w, J, dJdw)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nsteps, nsnap, i0, i1
    real(c_double), intent(in) :: wall_w(:,:)
    real(c_double), intent(out) :: J
    real(c_double), intent(out) :: dJdw(:,:)
    type(tbl_adjoint_opts) :: o
    integer(c_int) :: nf, na
    o%nsteps = nsteps; o%nsnap = nsnap; o%i0 = i0; o%i1 = i1; o%forward_only = 0; o%branch_mode = 0
    call gpu_check(ctx, c_adjoint_window(ctx, o, wall_w, J, dJdw, nf, na), 'tbl_adjoint_window')
  end subroutine adjoint_window

end module tbl_gpu
