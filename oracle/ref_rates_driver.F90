! Driver that runs the reference's own rate code (CALCULATE_REACTION_RATES
! with READ_SPECIES / READ_RATES) on points read from stdin and prints the
! rate coefficients.  Written for this repository as a checker of the C
! restatement in rates.c; it links object files compiled unchanged from the
! reference sources (see Makefile target "ref").  Run it with the working
! directory holding species_*.d / rates_*.d, as the reference reads them
! from the current directory (read_species.F90:91, read_rates.F90:45).
!
! stdin:  nspec nreac
!         zeta(normalised) v_turb(cm/s) omega grain_radius metallicity g2d
!         npts nrays
!         per point: T Td, G(1:nrays), Av(1:nrays), then for each ray the
!                    columns of H2 HD CO C S
! stdout: per point nreac rates, one per line (ES24.16E3)
program ref_rates_driver
  use definitions
  use healpix_types
  use global_module
  use maincode_module, only : zeta, v_turb
  use uclpdr_module, only : start, SCO_GRID
  implicit none
  integer(kind=i4b) :: nspec, nreac, npts, nrays, p, k, ic
  integer(kind=i4b) :: nrgr, nrh2, nrhd, nrco, nrci, nrsi
  real(kind=dp) :: T, Td
  real(kind=dp), allocatable :: G(:), Av(:), col5(:,:), column(:,:)
  real(kind=dp), allocatable :: abundance(:), mass(:), alpha(:), beta(:), gamma(:)
  real(kind=dp), allocatable :: rate(:), rtmin(:), rtmax(:)
  integer(kind=i4b), allocatable :: duplicate(:)
  character(len=10), allocatable :: species(:), reactant(:,:), product(:,:)
  integer(kind=i4b) :: idx(5)

  read(5,*) nspec, nreac
  read(5,*) zeta, v_turb, omega, grain_radius, metallicity, g2d
  read(5,*) npts, nrays
  allocate(G(0:nrays-1), Av(0:nrays-1), col5(0:nrays-1,5), column(0:nrays-1,1:nspec))
  allocate(abundance(nspec), mass(nspec), species(nspec))
  allocate(alpha(nreac), beta(nreac), gamma(nreac), rate(nreac), rtmin(nreac), rtmax(nreac))
  allocate(duplicate(nreac), reactant(nreac,3), product(nreac,4))

  open(10, file='/dev/null')
  call READ_SPECIES(nspec, species, abundance, mass)
  call READ_RATES(nreac, reactant, product, alpha, beta, gamma, rate, duplicate, rtmin, rtmax)

  ! CO shielding table (same values as the reference's allocations routine)
  SCO_GRID(1:8,1) = (/0.000D+00,-1.408D-02,-1.099D-01,-4.400D-01,-1.154D+00,-1.888D+00,-2.760D+00,-4.001D+00/)
  SCO_GRID(1:8,2) = (/-8.539D-02,-1.015D-01,-2.104D-01,-5.608D-01,-1.272D+00,-1.973D+00,-2.818D+00,-4.055D+00/)
  SCO_GRID(1:8,3) = (/-1.451D-01,-1.612D-01,-2.708D-01,-6.273D-01,-1.355D+00,-2.057D+00,-2.902D+00,-4.122D+00/)
  SCO_GRID(1:8,4) = (/-4.559D-01,-4.666D-01,-5.432D-01,-8.665D-01,-1.602D+00,-2.303D+00,-3.146D+00,-4.421D+00/)
  SCO_GRID(1:8,5) = (/-1.303D+00,-1.312D+00,-1.367D+00,-1.676D+00,-2.305D+00,-3.034D+00,-3.758D+00,-5.077D+00/)
  SCO_GRID(1:8,6) = (/-3.883D+00,-3.888D+00,-3.936D+00,-4.197D+00,-4.739D+00,-5.165D+00,-5.441D+00,-6.446D+00/)
  start = .true.

  idx = (/ NH2, NHD, NCO, NC, NS /)
  do p = 1, npts
    read(5,*) T, Td, (G(k), k=0,nrays-1), (Av(k), k=0,nrays-1), ((col5(k,ic), ic=1,5), k=0,nrays-1)
    column = 0.0D0
    do k = 0, nrays-1
      do ic = 1, 5
        if (idx(ic) > 0) column(k, idx(ic)) = col5(k, ic)
      end do
    end do
    call CALCULATE_REACTION_RATES(T, Td, nrays, G, Av, column, nreac, reactant, product, &
         & alpha, beta, gamma, rate, rtmin, rtmax, duplicate, nspec, nrgr, nrh2, nrhd, nrco, nrci, nrsi)
    do k = 1, nreac
      write(6, '(ES24.16E3)') rate(k)
    end do
  end do
end program ref_rates_driver
