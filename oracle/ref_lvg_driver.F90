! Driver that runs the reference's own level-population routines -- READINPUT
! (read_input.F90), FIND_CCOEFF (find_Ccoeff.F90), escape_probability
! (escape_probability.F90), solvlevpop (solvlevpop.F90) and
! CALCULATE_DUST_TEMPERATURES (dust_t.F90) -- on cases read from stdin, and
! prints their outputs.  Written for this repository as a checker of the C
! restatement in thermal.c; it links object files compiled unchanged from
! the reference sources (Makefile target "ref"), once with PSEUDO_1D (only
! ray 6 carries optical depth, as the reference's 1D build) and once without.
!
! The collision-rate arrays are zeroed before READINPUT is called (READINPUT
! zeroes all of them except the H+ partner's, read_input.F90:35-49; a fresh
! allocation holds zeros in practice).
!
! stdin, first line: mode
!   mode 1 (tables):  file (a line of its own), then nlev ntemp
!     stdout: energies, weights, A, B, frequencies (column-major), the
!     temperature table (7 x ntemp), the 7 partners' rates (H2, pH2, oH2, e,
!     H, He, H+, each nlev x nlev x ntemp column-major), one value per line
!   mode 2 (level populations): file (a line of its own), then nlev ntemp
!     coolant(1..4) ncases, then per case
!       T Td rho metallicity v_turb(cm/s) g2d nrays
!       nH nHp ne nHe nH2 (partner abundances, fractional) xspec (coolant's abundance)
!       pop(1:nlev)
!       per ray j: epray(j), then for i = 0..epray(j): x y z pop(1:nlev)
!     stdout per case: C (nlev^2), transition (nlev^2), cooling, line (nlev^2),
!     tau (nlev^2 x nrays), beta (nlev^2 x nrays), solution (nlev)
!   mode 3 (dust temperature): npts nrays av_fac, then per point UVfield and
!     rad_surface(0:nrays-1); stdout: dust temperature per point
program ref_lvg_driver
  use definitions
  use healpix_types
  use global_module, only : g2d
  use maincode_module, only : pdr, pdr_ptot, IDlist_pdr, nrays, AV_fac
  implicit none
  integer(kind=i4b) :: mode, nlev, ntemp, coolant, ncases, ic, j, i, k, maxpoints, npts
  character(len=256) :: fname
  real(kind=dp), allocatable :: E(:), W(:), A(:,:), B(:,:), F(:,:), TT(:,:)
  real(kind=dp), allocatable :: CH(:,:,:), CHP(:,:,:), CEL(:,:,:), CHE(:,:,:), CH2(:,:,:), CPH2(:,:,:), COH2(:,:,:)
  real(kind=dp), allocatable :: CC(:,:), TR(:,:), LN(:,:), TAU(:,:,:), BETA(:,:,:), SOL(:), POP(:)
  real(kind=dp), allocatable :: EVP(:,:,:), EPT(:,:,:)
  integer(kind=i4b), allocatable :: EPR(:)
  real(kind=dp) :: T, Td, rho, metal, vturb, xh, xhp, xe, xhe, xh2, xsp, cool
  integer(kind=i4b), parameter :: MAXP = 2000

  read(5,*) mode
  if (mode == 3) then
    read(5,*) npts, nrays, AV_fac
    pdr_ptot = npts
    allocate(pdr(1:npts), IDlist_pdr(1:npts))
    do i = 1, npts
      IDlist_pdr(i) = i
      allocate(pdr(i)%rad_surface(0:nrays-1))
      read(5,*) pdr(i)%UVfield, (pdr(i)%rad_surface(j), j=0,nrays-1)
    end do
    call CALCULATE_DUST_TEMPERATURES
    do i = 1, npts
      write(6, '(ES24.16E3)') pdr(i)%dust_t
    end do
    stop
  end if

  read(5,'(A)') fname
  if (mode == 1) then
    read(5,*) nlev, ntemp
  else
    read(5,*) nlev, ntemp, coolant, ncases
  end if
  allocate(E(nlev), W(nlev), A(nlev,nlev), B(nlev,nlev), F(nlev,nlev), TT(7,ntemp))
  allocate(CH(nlev,nlev,ntemp), CHP(nlev,nlev,ntemp), CEL(nlev,nlev,ntemp), CHE(nlev,nlev,ntemp))
  allocate(CH2(nlev,nlev,ntemp), CPH2(nlev,nlev,ntemp), COH2(nlev,nlev,ntemp))
  CH = 0.0D0; CHP = 0.0D0; CEL = 0.0D0; CHE = 0.0D0; CH2 = 0.0D0; CPH2 = 0.0D0; COH2 = 0.0D0
  call READINPUT(trim(fname), nlev, ntemp, E, W, A, B, F, TT, CH, CHP, CEL, CHE, CH2, CPH2, COH2)
  if (mode == 1) then
    call put(E, nlev); call put(W, nlev)
    call put(A, nlev*nlev); call put(B, nlev*nlev); call put(F, nlev*nlev)
    call put(TT, 7*ntemp)
    call put(CH2, nlev*nlev*ntemp); call put(CPH2, nlev*nlev*ntemp); call put(COH2, nlev*nlev*ntemp)
    call put(CEL, nlev*nlev*ntemp); call put(CH, nlev*nlev*ntemp); call put(CHE, nlev*nlev*ntemp)
    call put(CHP, nlev*nlev*ntemp)
    stop
  end if

  allocate(CC(nlev,nlev), TR(nlev,nlev), LN(nlev,nlev), SOL(nlev), POP(nlev))
  do ic = 1, ncases
    read(5,*) T, Td, rho, metal, vturb, g2d, nrays
    read(5,*) xh, xhp, xe, xhe, xh2, xsp
    read(5,*) (POP(k), k=1,nlev)
    if (allocated(EPR)) deallocate(EPR, EVP, EPT, TAU, BETA)
    allocate(EPR(0:nrays-1), EVP(0:nrays-1,0:MAXP,1:nlev), EPT(1:3,0:nrays-1,0:MAXP))
    allocate(TAU(nlev,nlev,0:nrays-1), BETA(nlev,nlev,0:nrays-1))
    EVP = 0.0D0; EPT = 0.0D0; TAU = 0.0D0; BETA = 0.0D0
    maxpoints = MAXP
    do j = 0, nrays-1
      read(5,*) EPR(j)
      do i = 0, EPR(j)
        read(5,*) EPT(1,j,i), EPT(2,j,i), EPT(3,j,i), (EVP(j,i,k), k=1,nlev)
      end do
    end do
    call FIND_CCOEFF(ntemp, nlev, T, TT, CH, CHP, CEL, CHE, CH2, CPH2, COH2, CC, &
         & xh*rho, xhp*rho, xe*rho, xhe*rho, xh2*rho, coolant)
    TR = 0.0D0
    call escape_probability(TR, Td, nrays, nlev, A, B, CC, F, EVP, maxpoints, T, vturb, EPR, POP, EPT, W, &
         & cool, LN, TAU, coolant, rho, metal, BETA)
    call put(CC, nlev*nlev)
    call put(TR, nlev*nlev)
    write(6, '(ES24.16E3)') cool
    call put(LN, nlev*nlev)
    call put(TAU, nlev*nlev*nrays)
    call put(BETA, nlev*nlev*nrays)
    call solvlevpop(nlev, TR, xsp*rho, SOL)
    call put(SOL, nlev)
  end do

contains

  subroutine put(x, n)
    integer(kind=i4b), intent(in) :: n
    real(kind=dp), intent(in) :: x(n)
    integer(kind=i4b) :: q
    do q = 1, n
      write(6, '(ES24.16E3)') x(q)
    end do
  end subroutine put

end program ref_lvg_driver
