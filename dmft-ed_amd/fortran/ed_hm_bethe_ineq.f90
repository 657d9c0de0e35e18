!> DMFT with inequivalent impurity sites: Nlat sites of the Bethe lattice
!> (semicircular DOS of half-bandwidth wband), site i with its own
!> interaction U_i = Uloc(1) + (i-1)*dU and bath, each closed by the
!> site-diagonal Bethe self-consistency.  It uses the lattice forms of the
!> DMFT_ED API, as the reference's inequivalent-site drivers do
!> (ed_init_solver(bath(Nlat,:),Hloc), ed_solve(bath,Hloc,Uloc_ii),
!> ed_get_sigma_matsubara(Smats,Nlat), ed_chi2_fitgf(bath,Weiss,Hloc,ispin);
!> ED_MAIN.f90:170-195, 364-468; ED_CHI2FIT.f90:413-452).  All sites of one
!> iteration are solved concurrently on the GPU.
!> Usage: ed_hm_bethe_ineq [input file] (default inputED.in); extra keys:
!> NLAT (2), DU (0.5), WBAND (1.0), LE (500), WMIXING (0.5).
program ed_hm_bethe_ineq
  use DMFT_ED
  implicit none
  integer                  :: iloop, Nb, Le, Nlat, i, ie, ilat, unit
  real(8)                  :: wband, wmixing, dU, de, pi
  real(8),allocatable      :: bath(:,:), bath_(:,:), He(:), Wte(:), wm(:), Uii(:,:)
  real(8),allocatable      :: dens(:,:), docc(:,:), mag(:,:), eimp(:,:), dbl(:,:)
  complex(8),allocatable   :: Hloc(:,:,:,:,:), Weiss(:,:,:,:,:,:), Smats(:,:,:,:,:,:)
  complex(8),allocatable   :: Gloc(:,:,:,:,:,:)
  complex(8),allocatable   :: S11(:,:,:,:,:), S21(:,:,:), S31(:), Gr(:,:,:,:,:,:), G11(:,:,:,:,:)
  character(len=256)       :: finput

  pi = acos(-1d0)
  finput = 'inputED.in'
  if(command_argument_count()>=1)call get_command_argument(1,finput)
  call ed_read_input(trim(finput))
  Nlat = 2; dU = 0.5d0; wband = 1d0; Le = 500; wmixing = 0.5d0
  call read_extra(trim(finput))

  allocate(He(Le), Wte(Le), wm(Lmats), Uii(Nlat,Norb))
  de = 2d0*wband/dble(Le-1)
  do ie=1,Le
     He(ie) = -wband + (ie-1)*de
     Wte(ie) = 2d0/(pi*wband**2)*sqrt(max(0d0,wband**2-He(ie)**2))*de
  enddo
  do i=1,Lmats
     wm(i) = pi/beta*dble(2*i-1)
  enddo
  do ilat=1,Nlat
     Uii(ilat,:) = Uloc(1:Norb) + (ilat-1)*dU
  enddo

  allocate(Weiss(Nlat,Nspin,Nspin,Norb,Norb,Lmats), Smats(Nlat,Nspin,Nspin,Norb,Norb,Lmats))
  allocate(Gloc(Nlat,Nspin,Nspin,Norb,Norb,Lmats))
  Weiss = (0d0,0d0); Gloc = (0d0,0d0)
  allocate(Hloc(Nlat,Nspin,Nspin,Norb,Norb)); Hloc = (0d0,0d0)
  Nb = get_bath_dimension()
  allocate(bath(Nlat,Nb), bath_(Nlat,Nb))
  allocate(dens(Nlat,Norb), docc(Nlat,Norb), mag(Nlat,Norb), eimp(Nlat,4), dbl(Nlat,4))
  call ed_init_solver(bath,Hloc)

  do iloop=1,nloop
     write(*,"(A,I4)")"DMFT-loop ",iloop
     call ed_solve(bath,Hloc,Uloc_ii=Uii)
     call ed_get_sigma_matsubara(Smats,Nlat)
     call bethe_gloc_weiss()
     call ed_chi2_fitgf(bath,Weiss,Hloc,ispin=1)
     if(iloop>1)bath = wmixing*bath + (1d0-wmixing)*bath_
     bath_ = bath
  enddo
  do ilat=1,Nlat
     open(newunit=unit,file="Gloc_iw_ineq"//itoa4(ilat)//".ed")
     do i=1,Lmats
        write(unit,"(3ES24.15)")wm(i),aimag(Gloc(ilat,1,1,1,1,i)),dble(Gloc(ilat,1,1,1,1,i))
     enddo
     close(unit)
  enddo
  call ed_get_dens(dens,Nlat)
  call ed_get_docc(docc,Nlat)
  call ed_get_mag(mag,Nlat)
  call ed_get_eimp(eimp,Nlat)
  call ed_get_doubles(dbl,Nlat)
  open(newunit=unit,file="dmft_summary_ineq.ed")
  do ilat=1,Nlat
     write(unit,"(A,I4,100ES24.15)")"site ",ilat,dens(ilat,:),docc(ilat,:),mag(ilat,:),&
          eimp(ilat,:),dbl(ilat,:)
  enddo
  close(unit)
  ! the one-site getter forms (_11, _21, _31) against the all-sites form
  allocate(S11(Nspin,Nspin,Norb,Norb,Lmats), S21(Nspin*Norb,Nspin*Norb,Lmats), S31(Lmats))
  allocate(Gr(Nlat,Nspin,Nspin,Norb,Norb,Lreal), G11(Nspin,Nspin,Norb,Norb,Lreal))
  call ed_get_sigma_matsubara(Smats,Nlat)
  call ed_get_sigma_matsubara(S11,Nlat)
  call ed_get_sigma_matsubara(S21,Nlat)
  call ed_get_sigma_matsubara(S31,Nlat,1,1,1,1)
  call ed_get_gimp_real(Gr,Nlat)
  call ed_get_gimp_real(G11,2)
  open(newunit=unit,file="getters_ineq.ed")
  write(unit,"(A,5ES24.15)")"getters ",maxval(abs(S11-Smats(Nlat,:,:,:,:,:))),&
       maxval(abs(S21(1,1,:)-Smats(Nlat,1,1,1,1,:))),maxval(abs(S31-Smats(Nlat,1,1,1,1,:))),&
       maxval(abs(G11-Gr(2,:,:,:,:,:))),maxval(abs(Smats(Nlat,1,1,1,1,:)-Smats(1,1,1,1,1,:)))
  close(unit)

contains

  subroutine bethe_gloc_weiss()
    integer    :: l, s, a, n, k
    complex(8) :: zeta, g
    do l=1,Nlat
       do s=1,Nspin
          do a=1,Norb
             do n=1,Lmats
                zeta = cmplx(xmu,wm(n),8) - Smats(l,s,s,a,a,n)
                g = (0d0,0d0)
                do k=1,Le
                   g = g + Wte(k)/(zeta - He(k))
                enddo
                Gloc(l,s,s,a,a,n) = g
                if(trim(cg_scheme)=='weiss')then
                   Weiss(l,s,s,a,a,n) = 1d0/(1d0/g + Smats(l,s,s,a,a,n))
                else
                   Weiss(l,s,s,a,a,n) = zeta - 1d0/g
                endif
             enddo
          enddo
       enddo
    enddo
  end subroutine bethe_gloc_weiss

  function itoa4(i) result(s)
    integer :: i
    character(len=4) :: s
    write(s,"(I4.4)")i
  end function itoa4

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
       case("NLAT","Nlat","nlat"); read(line(ieq+1:),*,iostat=ios)Nlat
       case("DU","dU","du");       read(line(ieq+1:),*,iostat=ios)dU
       case("WBAND","wband");      read(line(ieq+1:),*,iostat=ios)wband
       case("LE","Le","le");       read(line(ieq+1:),*,iostat=ios)Le
       case("WMIXING","wmixing");  read(line(ieq+1:),*,iostat=ios)wmixing
       end select
    enddo
    close(u)
  end subroutine read_extra

end program ed_hm_bethe_ineq
