! Writes one record of each FORMAT statement of the reference's file formats
! (dnssetbl/dnssetbl.f90 formats 101-115, dnssetbl/combineFile.f90 105/106)
! for a fixed set of values; tests/test_refio.py compares the bytes with
! turbulentbl_amd/refio.py.
program format_probe
  implicit none
  real(8) :: v(7), z, nan, inf
  v = [0.5d0, 12.0d0, 0.0013d0, 19.798123456789012d0, -0.000123456789d0, 0d0, 1.5d-120]
  z = 0d0
  nan = z / z
  inf = 1d0 / z
  write(*,104) v
  write(*,104) 11.9882697947214d0, 3.96825396825397d0, 3.4d0, -1.0d-300, 2.5d300, 1.0d0, -9.99999999999999999d-1
  write(*,105) 'TITLE ="', 1.2345d0, 1500, '"'
  write(*,107) 'TITLE ="', 1.2345d0, 1500, 1, '"'
  write(*,106) 'ZONE k=', 1024, ',j=', 256, ',i=', 192, ',DATAPACKING=POINT'
  write(*,108) 'ZONE k=', 1024, ',j=', 254, ',i=', 192, ',DATAPACKING="POINT"'
  write(*,115) 19.5d0, -0.25d0, 0.125d0, 0.99d0, 7
  write(*,115) 0.0d0, 1.0d-8, -19.798d0, 1.0d0, 123
  write(*,'(24x,E18.8)') 19.798d0
  ! statistics files of the rescale (dnssetbl.f90:1071-1198)
  write(*,110) 0.1d0, 1.0000000001d0, -2d0
  write(*,111) -1.0d-8, nan, inf
  write(*,111) -inf, 1d9, -1.5d0
  write(*,112) 1d0, 0.0123456789012d0, 2.5d0, 3d0, 4d0, 500.25d0, 0d0
  write(*,113) 7, 12, 1.25d0
  write(*,114) 1d0, 2d0, 3d0, 4d0, 5d0
  write(*,116) 1d0, 2d0, 3d0, 4d0, 5d0, 6d0, 7d0, -8d0
  write(*,*) 'Variables = "x1","x3","Wt"'
110 format(3(2x,F20.10))
111 format(3(1X,F15.7))
112 format(8(2x,F20.10))
113 format(2(I5),F20.10)
114 format(5(1X,F15.7))
116 format(8(1X,F15.7))
104 FORMAT(3(1X,F10.6),4(1X,E22.15))
105 format(A,F20.10,I9,A)
106 format(A,I4,A,I4,A,I4,A)
107 FORMAT(A,F20.10,I9,I3,A)
108 FORMAT(A,I4,A,I4,A,I4,A)
115 FORMAT(4(1X,F15.7),I5.3)
end program
