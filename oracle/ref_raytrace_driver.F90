! Driver that runs the reference's own ray tracer (evaluation_points,
! eval_points.F90, with the HEALPix routines of healpix.F90 and heapsort.F90)
! on a point cloud read from stdin and writes the evaluation points it finds.
! Written for this repository as a checker of the C restatement in
! raytrace.c; it links object files compiled unchanged from the reference
! sources (see Makefile target "ref").
!
! stdin:  npts level theta_crit maxpoints
!         per point: x y z etype
! file raytrace_out.txt (in the working directory):
!         nrays lines "vx vy vz" (HEALPix ray vectors, pix2vec_nest)
!         per point p and ray j: "p j epray", then epray lines
!         "projected x y z" (0-based point ids, evaluation point coordinates)
program ref_raytrace_driver
  use definitions
  use healpix_types
  use healpix_module
  use maincode_module
  implicit none
  integer(kind=i4b) :: np, q, jr, ie
  real(kind=dp) :: vec(1:3), vtx(1:3,1:4)

  read(5,*) np, level, theta_crit, maxpoints
  nside = 2**level
  nrays = 12*nside**2
  grand_ptot = np
  pdr_ptot = np
  dark_ptot = 0
  pix2x = 0
  pix2y = 0
  x2pix = 0
  y2pix = 0
  allocate(vectors(1:3,0:nrays-1))
  do jr = 0, nrays-1
     call pix2vec_nest(nside, jr, pix2x, pix2y, vec, vtx)
     vectors(1:3,jr) = vec(1:3)
  enddo
  allocate(pdr(1:np), pdrpoint(1:3,1:np), IDlist_pdr(1:np), IDlist_dark(1:1))
  do q = 1, np
     read(5,*) pdr(q)%x, pdr(q)%y, pdr(q)%z, pdr(q)%etype
     pdrpoint(1,q) = pdr(q)%x
     pdrpoint(2,q) = pdr(q)%y
     pdrpoint(3,q) = pdr(q)%z
     IDlist_pdr(q) = q
     allocate(pdr(q)%epray(0:nrays-1))
     allocate(pdr(q)%epoint(1:3,0:nrays-1,0:maxpoints))
     allocate(pdr(q)%projected(0:nrays-1,0:maxpoints))
  enddo

  call evaluation_points

  open(20, file='raytrace_out.txt', status='replace')
  do jr = 0, nrays-1
     write(20,'(3ES26.17E3)') vectors(1:3,jr)
  enddo
  do q = 1, np
     do jr = 0, nrays-1
        write(20,'(3I10)') q-1, jr, pdr(q)%epray(jr)
        do ie = 1, pdr(q)%epray(jr)
           write(20,'(I10,3ES26.17E3)') pdr(q)%projected(jr,ie)-1, pdr(q)%epoint(1:3,jr,ie)
        enddo
     enddo
  enddo
  close(20)
end program ref_raytrace_driver
