! A stand-in for the 3D-PDR main program at the chemistry call: it uses the
! reference's own modules (modules.F90, compiled unchanged from the reference
! sources by the Makefile target "ref"), so the tolerances, start/end time and
! electron index live in the Fortran modules' bind(c) variables exactly as in
! 3D-PDR (modules.F90:165-166, 383, 529-530), and it calls the chemistry with
! the reference's implicit interface (3DPDR.F90:498).  Linked against
! libpdrchem.so instead of calculate_abundances.o + SUNDIALS, it shows the
! drop-in boundary from Fortran.  Written for this repository.
!
! Arguments: input file, output file (unformatted stream):
!   in:  npart nspec nreac nelect (int32), rtol atol start_time end_time (real64),
!        abundance(nspec,npart), rate(nreac,npart), density(npart), temperature(npart)
!   out: status (int32), abundance(nspec,npart)
program ref_host_driver
  use definitions
  use healpix_types
  use chemistry_module, only : relative_abundance_tolerance, absolute_abundance_tolerance
  use maincode_module, only : start_time, end_time
  use global_module, only : NELECT
  implicit none
  integer(kind=i4b) :: npart, nspec, nreac, ne, status
  real(kind=dp) :: rtol, atol, t0, t1
  real(kind=dp), allocatable :: abundance(:,:), rate(:,:), density(:), temperature(:)
  character(len=1024) :: fin, fout
  integer(kind=i4b), external :: calculate_abundances

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) npart, nspec, nreac, ne
  read(10) rtol, atol, t0, t1
  allocate(abundance(nspec, npart), rate(nreac, npart), density(npart), temperature(npart))
  read(10) abundance, rate, density, temperature
  close(10)

  relative_abundance_tolerance = rtol
  absolute_abundance_tolerance = atol
  start_time = t0
  end_time = t1
  NELECT = ne

  status = CALCULATE_ABUNDANCES(abundance, rate, density, temperature, npart, nspec, nreac)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) status
  write(11) abundance
  close(11)
end program ref_host_driver
