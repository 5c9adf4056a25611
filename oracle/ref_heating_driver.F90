! Driver that runs the reference's own heating routine (CALC_HEATING,
! sub_calculate_heating.F90) on points read from stdin and prints the 12
! heating terms.  Written for this repository as a checker of the C
! restatement in heating.c; it links object files compiled unchanged from the
! reference sources (Makefile target "ref").  Species and reaction indices come
! from the reference's READ_SPECIES and CALCULATE_REACTION_RATES; run it with
! the working directory holding species_*.d / rates_*.d.
!
! stdin:  nspec nreac
!         zeta(normalised) v_turb(cm/s) omega grain_radius metallicity g2d
!         npts
!         per point: n_H T Td UVfield, then nspec abundances, then nreac rates
! stdout: per point 12 heating terms, one per line (ES24.16E3)
program ref_heating_driver
  use definitions
  use healpix_types
  use global_module
  use maincode_module, only : zeta, v_turb
  use uclpdr_module, only : start
  implicit none
  integer(kind=i4b) :: nspec, nreac, npts, p, k
  integer(kind=i4b) :: nrgr, nrh2, nrhd, nrco, nrci, nrsi
  real(kind=dp) :: dens, T, Td, uv, heat(12), G(0:0), Av(0:0)
  real(kind=dp), allocatable :: column(:,:), abundance(:), mass(:), alpha(:), beta(:), gamma(:)
  real(kind=dp), allocatable :: rate(:), rtmin(:), rtmax(:)
  integer(kind=i4b), allocatable :: duplicate(:)
  character(len=10), allocatable :: species(:), reactant(:,:), product(:,:)

  read(5,*) nspec, nreac
  read(5,*) zeta, v_turb, omega, grain_radius, metallicity, g2d
  read(5,*) npts
  allocate(column(0:0,1:nspec), abundance(nspec), mass(nspec), species(nspec))
  allocate(alpha(nreac), beta(nreac), gamma(nreac), rate(nreac), rtmin(nreac), rtmax(nreac))
  allocate(duplicate(nreac), reactant(nreac,3), product(nreac,4))
  open(10, file='/dev/null')
  call READ_SPECIES(nspec, species, abundance, mass)
  call READ_RATES(nreac, reactant, product, alpha, beta, gamma, rate, duplicate, rtmin, rtmax)
  start = .true.
  ! the reaction indices NRGR, NRH2, NRCI the heating reads
  G = 0.0D0
  Av = 0.0D0
  column = 0.0D0
  call CALCULATE_REACTION_RATES(10.0D0, 10.0D0, 1, G, Av, column, nreac, reactant, product, &
       & alpha, beta, gamma, rate, rtmin, rtmax, duplicate, nspec, nrgr, nrh2, nrhd, nrco, nrci, nrsi)
  do p = 1, npts
    read(5,*) dens, T, Td, uv
    read(5,*) (abundance(k), k=1,nspec)
    read(5,*) (rate(k), k=1,nreac)
    call CALC_HEATING(dens, T, Td, uv, v_turb, nspec, abundance, nreac, rate, heat, &
         & nrgr, nrh2, nrhd, nrco, nrci, nrsi)
    do k = 1, 12
      write(6, '(ES24.16E3)') heat(k)
    end do
  end do
end program ref_heating_driver
