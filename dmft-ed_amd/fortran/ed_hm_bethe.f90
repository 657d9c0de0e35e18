!> DMFT for the Hubbard model on the Bethe lattice (semicircular DOS of
!> half-bandwidth wband), the loop of the reference driver
!> drivers/ed_hm_bethe.f90:20-91, written against the DMFT_ED API only:
!>   ed_solve -> Sigma(iw) -> G_loc(iw) = sum_e rho(e) de / (iw + xmu - e - Sigma)
!>   -> Weiss field 1/(1/G_loc + Sigma) (cg_scheme weiss) or
!>      Delta = iw + xmu - Sigma - 1/G_loc (cg_scheme delta)
!>   -> ed_chi2_fitgf -> linear mixing of the bath.
!> The DOS is discretised on Le points of linspace(-wband, wband) as the
!> reference does (:38-40).  Usage: ed_hm_bethe [input file] (default inputED.in);
!> extra keys in the input: WBAND (1.0), LE (500), WMIXING (0.5).
program ed_hm_bethe
  use DMFT_ED
  implicit none
  integer                  :: iloop, Nb, Le, i, ie, nsucc, unit
  logical                  :: converged
  real(8)                  :: wband, wmixing, de, err, pi, eimp(4), dbl(4)
  real(8),allocatable      :: mag(:)
  real(8),allocatable      :: bath(:), bath_(:), He(:), Wte(:), wm(:)
  complex(8),allocatable   :: Hloc(:,:,:,:), Weiss(:,:,:,:,:), Weiss_old(:,:,:,:,:)
  complex(8),allocatable   :: Smats(:,:,:,:,:), Gloc(:,:,:,:,:), S2(:,:,:), S3(:)
  real(8)                  :: d2
  character(len=256)       :: finput

  pi = acos(-1d0)
  finput = 'inputED.in'
  if(command_argument_count()>=1)call get_command_argument(1,finput)
  call ed_read_input(trim(finput))
  wband = 1d0; Le = 500; wmixing = 0.5d0
  call read_extra(trim(finput))

  allocate(He(Le), Wte(Le), wm(Lmats), mag(Norb))
  de = 2d0*wband/dble(Le-1)
  do ie=1,Le
     He(ie) = -wband + (ie-1)*de
     Wte(ie) = 2d0/(pi*wband**2)*sqrt(max(0d0,wband**2-He(ie)**2))*de   ! dens_bethe * de
  enddo
  do i=1,Lmats
     wm(i) = pi/beta*dble(2*i-1)
  enddo

  allocate(Weiss(Nspin,Nspin,Norb,Norb,Lmats), Weiss_old(Nspin,Nspin,Norb,Norb,Lmats))
  allocate(Smats(Nspin,Nspin,Norb,Norb,Lmats), Gloc(Nspin,Nspin,Norb,Norb,Lmats))
  ! orbital-diagonal self-consistency: the off-diagonal Weiss elements stay 0
  Weiss = (0d0,0d0); Gloc = (0d0,0d0)
  allocate(Hloc(Nspin,Nspin,Norb,Norb)); Hloc = (0d0,0d0)
  Nb = get_bath_dimension()
  allocate(bath(Nb), bath_(Nb))
  call ed_init_solver(bath,Hloc)

  iloop = 0; converged = .false.; nsucc = 0
  Weiss_old = (0d0,0d0)
  do while(.not.converged .and. iloop<nloop)
     iloop = iloop + 1
     write(*,"(A,I4)")"DMFT-loop ",iloop
     call ed_solve(bath)
     call ed_get_sigma_matsubara(Smats)
     call bethe_gloc_weiss()
     call ed_chi2_fitgf(Weiss,bath,ispin=1)
     if(iloop>1)bath = wmixing*bath + (1d0-wmixing)*bath_
     bath_ = bath
     err = sum(abs(Weiss(1,1,1,1,:)-Weiss_old(1,1,1,1,:)))/sum(abs(Weiss(1,1,1,1,:)))
     Weiss_old = Weiss
     write(*,"(A,ES14.6)")"error=",err
     if(err<dmft_error)then
        nsucc = nsucc + 1
     else
        nsucc = 0
     endif
     converged = nsucc>=Nsuccess
  enddo
  open(newunit=unit,file="Gloc_iw.ed")
  do i=1,Lmats
     write(unit,"(3ES24.15)")wm(i),aimag(Gloc(1,1,1,1,i)),dble(Gloc(1,1,1,1,i))
  enddo
  close(unit)
  open(newunit=unit,file="dmft_summary.ed")
  write(unit,"(A,I6)")"iloops ",iloop
  write(unit,"(A,ES24.15)")"egs ",ed_egs
  write(unit,"(A,10ES24.15)")"dens ",ed_dens
  write(unit,"(A,10ES24.15)")"docc ",ed_docc
  call ed_get_eimp(eimp)
  call ed_get_doubles(dbl)
  call ed_get_mag(mag)
  write(unit,"(A,4ES24.15)")"eimp ",eimp
  write(unit,"(A,4ES24.15)")"doubles ",dbl
  write(unit,"(A,10ES24.15)")"mag ",mag
  ! the (Nspin*Norb)^2 and single-element getter forms agree with the full one
  call ed_get_sigma_matsubara(Smats)
  allocate(S2(Nspin*Norb,Nspin*Norb,Lmats), S3(Lmats))
  call ed_get_sigma_matsubara(S2)
  call ed_get_sigma_matsubara(S3,1,1,1,1)
  call ed_get_dens(d2,1)
  write(unit,"(A,3ES24.15)")"getters ",maxval(abs(S2(1,1,:)-Smats(1,1,1,1,:))),&
       maxval(abs(S3-Smats(1,1,1,1,:))),abs(d2-ed_dens(1))
  close(unit)

contains

  subroutine bethe_gloc_weiss()
    integer    :: s, a, n, k
    complex(8) :: zeta, g
    do s=1,Nspin
       do a=1,Norb
          do n=1,Lmats
             zeta = cmplx(xmu,wm(n),8) - Smats(s,s,a,a,n)
             g = (0d0,0d0)
             do k=1,Le
                g = g + Wte(k)/(zeta - He(k))
             enddo
             Gloc(s,s,a,a,n) = g
             if(trim(cg_scheme)=='weiss')then
                Weiss(s,s,a,a,n) = 1d0/(1d0/g + Smats(s,s,a,a,n))
             else
                Weiss(s,s,a,a,n) = zeta - 1d0/g
             endif
          enddo
       enddo
    enddo
  end subroutine bethe_gloc_weiss

  subroutine read_extra(fname)
    character(len=*) :: fname
    integer          :: u, ios, ieq, iex
    character(len=512) :: line, key
    logical :: ex
    inquire(file=fname,exist=ex)
    if(.not.ex)return
    open(newunit=u,file=fname,status='old',action='read')
    do
       read(u,"(A)",iostat=ios)line
       if(ios/=0)exit
       iex = index(line,'!'); if(iex>0)line=line(:iex-1)
       ieq = index(line,'='); if(ieq<=1)cycle
       key = adjustl(line(:ieq-1))
       select case(trim(key))
       case("WBAND","wband"); read(line(ieq+1:),*,iostat=ios)wband
       case("LE","Le","le");   read(line(ieq+1:),*,iostat=ios)Le
       case("WMIXING","wmixing"); read(line(ieq+1:),*,iostat=ios)wmixing
       end select
    enddo
    close(u)
  end subroutine read_extra

end program ed_hm_bethe
