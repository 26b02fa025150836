! prints the library's default parameters as seen through the bind(C) type
! of tbl_gpu: checks that the Fortran layout of struct tbl_params matches C
program layout_probe
  use iso_c_binding
  use tbl_gpu
  implicit none
  type(tbl_params), target :: p
  integer(c_intptr_t) :: base
  call gpu_default_params(p)
  write(*,'(3i6)') p%nX, p%nY, p%nZ
  write(*,'(es24.16)') p%omega, p%Re, p%Ue, p%delta, p%errtol, p%errorstop, p%inner_tol
  write(*,'(5i6)') p%planes
  write(*,'(9i6)') p%i_rec, p%freq_tim_avg, p%precision, p%gs_mmax, p%helm_mode, p%tol_mode, p%kmax, p%mmax, p%precond
  write(*,'(i6)') int(c_sizeof(p))
  ! byte offsets of the last members (scheme, forc)
  base = transfer(c_loc(p), base)
  write(*,'(2i6)') int(transfer(c_loc(p%scheme), base) - base), int(transfer(c_loc(p%forc), base) - base)
  write(*,'(i6, 3es24.16)') p%scheme, p%forc
  write(*,'(2i6)') int(transfer(c_loc(p%slabs), base) - base), p%slabs
end program layout_probe
