!> The reference's operator-level interface to H*v, on the GPU library.
!>
!> The reference applies H through the procedure pointer spHtimesV_cc of
!> abstract interface cc_sparse_HxV(Nloc,v,Hv) with complex(8) vectors
!> (ED_VARS_GLOBAL.f90:43-49,100), bound to directMatVec_cc in ed_solve_single
!> (ED_MAIN.f90:269-276), for the "current sector" selected by
!> setup_Hv_sector(isector) and released by delete_Hv_sector
!> (ED_HAMILTONIAN_MATVEC.f90:99-111).  This module keeps those names and
!> shapes, so a driver or an ED_DIAG that keeps the reference's own Lanczos
!> and only swaps the operator points spHtimesV_cc at directMatVec_cc here.
!>
!>   setup_Hv_sector(isector) : hands the model of module DMFT_ED (bath set by
!>                              set_bath/ed_solve, Hloc, U, ...) to the library
!>                              and builds sector isector = nup*(Ns+1)+ndw+1
!>                              with two device vectors for the product
!>   directMatVec_cc(Nloc,v,Hv): Hv = H v for the current sector.  H is real
!>                              (normal or hybrid bath, real Hloc), so the
!>                              real and imaginary parts go through one FP64
!>                              product each (one when v is real)
!>   delete_Hv_sector()       : releases the sector and its vectors
!>
!> Vector order is the reference's (i = r_up + r_dw*D_up); errors stop with
!> the library's message, as the reference stops.
module ED_HXV_GPU
  use iso_c_binding
  use DMED_CABI
  use DMFT_ED, only: Nspin, Norb, Nbath, bath_type, ed_push_model, ed_check
  implicit none
  private

  abstract interface
     subroutine cc_sparse_HxV(Nloc,v,Hv)
       integer                 :: Nloc
       complex(8),dimension(Nloc) :: v
       complex(8),dimension(Nloc) :: Hv
     end subroutine cc_sparse_HxV
  end interface

  public :: cc_sparse_HxV
  public :: setup_Hv_sector, delete_Hv_sector, directMatVec_cc, Hv_sector_dimension

  type(c_ptr),save :: hsec = c_null_ptr, xv = c_null_ptr, yv = c_null_ptr
  integer(c_int64_t),save :: hdim = 0

contains

  integer function sites_count()
    if(trim(bath_type)=='hybrid')then
       sites_count = Nbath+Norb
    else
       sites_count = (Nbath+1)*Norb
    endif
  end function sites_count

  subroutine setup_Hv_sector(isector)
    integer,intent(in) :: isector
    integer            :: Ns, nup, ndw
    integer(c_int64_t) :: du, dd
    if(c_associated(hsec))call delete_Hv_sector()
    Ns = sites_count()
    if(isector<1 .or. isector>(Ns+1)**2)stop "setup_Hv_sector error: isector out of range"
    nup = (isector-1)/(Ns+1)
    ndw = mod(isector-1,Ns+1)
    call ed_push_model()
    call ed_check(dmed_sector_create(int(nup,c_int32_t),int(ndw,c_int32_t),hsec),"dmed_sector_create")
    call ed_check(dmed_sector_dims(hsec,hdim,du,dd),"dmed_sector_dims")
    call ed_check(dmed_vec_create(hsec,DMED_FP64,xv),"dmed_vec_create")
    call ed_check(dmed_vec_create(hsec,DMED_FP64,yv),"dmed_vec_create")
  end subroutine setup_Hv_sector

  subroutine delete_Hv_sector()
    if(c_associated(xv))call ed_check(dmed_vec_free(xv),"dmed_vec_free")
    if(c_associated(yv))call ed_check(dmed_vec_free(yv),"dmed_vec_free")
    if(c_associated(hsec))call ed_check(dmed_sector_free(hsec),"dmed_sector_free")
    xv = c_null_ptr; yv = c_null_ptr; hsec = c_null_ptr
    hdim = 0
  end subroutine delete_Hv_sector

  integer function Hv_sector_dimension()
    Hv_sector_dimension = int(hdim)
  end function Hv_sector_dimension

  subroutine directMatVec_cc(Nloc,v,Hv)
    integer                    :: Nloc
    complex(8),dimension(Nloc) :: v
    complex(8),dimension(Nloc) :: Hv
    real(8),allocatable        :: xr(:), yr(:), yi(:)
    if(.not.c_associated(hsec))stop "directMatVec_cc error: no sector set (setup_Hv_sector)"
    if(int(Nloc,c_int64_t)/=hdim)stop "directMatVec_cc error: wrong dimension"
    allocate(xr(Nloc), yr(Nloc))
    xr = dble(v)
    call ed_check(dmed_vec_upload(xv,xr),"dmed_vec_upload")
    call ed_check(dmed_hxv(hsec,dmed_vec_data(xv),dmed_vec_data(yv),DMED_FP64),"dmed_hxv")
    call ed_check(dmed_vec_download(yv,yr),"dmed_vec_download")
    if(any(aimag(v)/=0d0))then
       allocate(yi(Nloc))
       xr = aimag(v)
       call ed_check(dmed_vec_upload(xv,xr),"dmed_vec_upload")
       call ed_check(dmed_hxv(hsec,dmed_vec_data(xv),dmed_vec_data(yv),DMED_FP64),"dmed_hxv")
       call ed_check(dmed_vec_download(yv,yi),"dmed_vec_download")
       Hv = cmplx(yr,yi,kind=8)
    else
       Hv = cmplx(yr,0d0,kind=8)
    endif
  end subroutine directMatVec_cc

end module ED_HXV_GPU
