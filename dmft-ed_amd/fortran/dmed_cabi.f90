!> ISO_C_BINDING interface to libdmed.so (include/dmed.h).
!> Derived types mirror the C structs field by field (bind(C) layout).
module DMED_CABI
  use iso_c_binding
  implicit none
  private

  integer(c_int),parameter,public :: DMED_FP64 = 0, DMED_FP32 = 1, DMED_MIXED = 2
  integer(c_int),parameter,public :: DMED_BATH_NORMAL = 0, DMED_BATH_HYBRID = 1

  type,bind(C),public :: dmed_model_t
     integer(c_int32_t) :: norb, nbath, nspin
     type(c_ptr)        :: bath, hloc
     real(c_double)     :: uloc(3)
     real(c_double)     :: ust, jh, jx, jp, xmu
     integer(c_int32_t) :: hfmode, jhflag
     integer(c_int32_t) :: bath_type = 0, pad_ = 0
  end type dmed_model_t

  type,bind(C),public :: dmed_solve_params_t
     integer(c_int32_t) :: lanc_nstates_sector, lanc_nstates_total, lanc_nstates_step
     integer(c_int32_t) :: lanc_niter, lanc_ngfiter, lanc_ncv_factor, lanc_ncv_add
     integer(c_int32_t) :: lanc_dim_threshold
     real(c_double)     :: lanc_tolerance, gs_threshold, cutoff, beta, eps, wini, wfin
     integer(c_int32_t) :: lmats, lreal, prec_gf, batched, verbose, gf_fp64_head, prec_gs
     integer(c_int32_t) :: chi_spin, ltau, chi_dens
  end type dmed_solve_params_t

  !> dmed_local_obs: orbital arrays (3), orbital pairs (9, [a*3+b] in C)
  type,bind(C),public :: dmed_local_obs_t
     real(c_double) :: dens(3), dens_up(3), dens_dw(3), docc(3), magz(3)
     real(c_double) :: sz2(9), n2(9)
     real(c_double) :: s2tot
     real(c_double) :: eknot, epot, ehartree, dust, dund, dse, dph, eint
  end type dmed_local_obs_t

  type,bind(C),public :: dmed_solve_out_t
     real(c_double)     :: egs, zeta
     integer(c_int32_t) :: nstates, numgs
     type(c_ptr)        :: dens, docc, gmats, greal, smats, sreal
     type(c_ptr)        :: neigen_sector, lanc_nstates_total
     real(c_double)     :: t_diag, t_gf
     type(c_ptr)        :: obs
     type(c_ptr)        :: spin_chi_iv, spin_chi_tau, spin_chi_w
     integer(c_int32_t) :: ltau_used, pad2_
     type(c_ptr)        :: dens_chi_iv, dens_chi_tau, dens_chi_w
     type(c_ptr)        :: dens_chi_mix_iv, dens_chi_mix_tau, dens_chi_mix_w
     type(c_ptr)        :: dens_chi_tot_iv, dens_chi_tot_tau, dens_chi_tot_w
  end type dmed_solve_out_t

  public :: dmed_init, dmed_finalize, dmed_last_error, dmed_set_model, dmed_solve
  public :: dmed_solve_sites
  public :: dmed_sector_create, dmed_sector_free, dmed_sector_dims, dmed_vec_create, dmed_vec_free
  public :: dmed_vec_upload, dmed_vec_download, dmed_vec_data, dmed_hxv
  public :: dmed_solve_params_default, dmed_sigma, dmed_delta, dmed_chi2_fit
  public :: dmed_chi2_fit_hybrid
  public :: c_string

  interface
     integer(c_int) function dmed_init(device) bind(C,name="dmed_init")
       import :: c_int
       integer(c_int),value :: device
     end function dmed_init
     integer(c_int) function dmed_finalize() bind(C,name="dmed_finalize")
       import :: c_int
     end function dmed_finalize
     type(c_ptr) function dmed_last_error() bind(C,name="dmed_last_error")
       import :: c_ptr
     end function dmed_last_error
     integer(c_int) function dmed_set_model(m) bind(C,name="dmed_set_model")
       import :: c_int, dmed_model_t
       type(dmed_model_t),intent(in) :: m
     end function dmed_set_model
     integer(c_int) function dmed_sector_create(nup,ndw,sec) bind(C,name="dmed_sector_create")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t),value :: nup, ndw
       type(c_ptr),intent(out)  :: sec
     end function dmed_sector_create
     integer(c_int) function dmed_sector_free(sec) bind(C,name="dmed_sector_free")
       import :: c_int, c_ptr
       type(c_ptr),value :: sec
     end function dmed_sector_free
     integer(c_int) function dmed_sector_dims(sec,dim,dimup,dimdw) bind(C,name="dmed_sector_dims")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr),value         :: sec
       integer(c_int64_t),intent(out) :: dim, dimup, dimdw
     end function dmed_sector_dims
     integer(c_int) function dmed_vec_create(sec,prec,v) bind(C,name="dmed_vec_create")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr),value        :: sec
       integer(c_int32_t),value :: prec
       type(c_ptr),intent(out)  :: v
     end function dmed_vec_create
     integer(c_int) function dmed_vec_free(v) bind(C,name="dmed_vec_free")
       import :: c_int, c_ptr
       type(c_ptr),value :: v
     end function dmed_vec_free
     integer(c_int) function dmed_vec_upload(v,x) bind(C,name="dmed_vec_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr),value         :: v
       real(c_double),intent(in) :: x(*)
     end function dmed_vec_upload
     integer(c_int) function dmed_vec_download(v,x) bind(C,name="dmed_vec_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr),value :: v
       real(c_double)    :: x(*)
     end function dmed_vec_download
     type(c_ptr) function dmed_vec_data(v) bind(C,name="dmed_vec_data")
       import :: c_ptr
       type(c_ptr),value :: v
     end function dmed_vec_data
     integer(c_int) function dmed_hxv(sec,x,y,prec) bind(C,name="dmed_hxv")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr),value        :: sec, x, y
       integer(c_int32_t),value :: prec
     end function dmed_hxv
     subroutine dmed_solve_params_default(p) bind(C,name="dmed_solve_params_default")
       import :: dmed_solve_params_t
       type(dmed_solve_params_t),intent(out) :: p
     end subroutine dmed_solve_params_default
     integer(c_int) function dmed_solve(p,o) bind(C,name="dmed_solve")
       import :: c_int, dmed_solve_params_t, dmed_solve_out_t
       type(dmed_solve_params_t),intent(in) :: p
       type(dmed_solve_out_t),intent(inout) :: o
     end function dmed_solve
     integer(c_int) function dmed_solve_sites(nsites,models,p,outs) bind(C,name="dmed_solve_sites")
       import :: c_int, c_int32_t, dmed_model_t, dmed_solve_params_t, dmed_solve_out_t
       integer(c_int32_t),value             :: nsites
       type(dmed_model_t),intent(in)        :: models(*)
       type(dmed_solve_params_t),intent(in) :: p
       type(dmed_solve_out_t),intent(inout) :: outs(*)
     end function dmed_solve_sites
     integer(c_int) function dmed_sigma(ispin,iorb,z,g,L,sigma) bind(C,name="dmed_sigma")
       import :: c_int, c_int32_t, c_double_complex
       integer(c_int32_t),value :: ispin, iorb, L
       complex(c_double_complex),intent(in) :: z(*), g(*)
       complex(c_double_complex),intent(out) :: sigma(*)
     end function dmed_sigma
     integer(c_int) function dmed_delta(ispin,iorb,z,L,delta) bind(C,name="dmed_delta")
       import :: c_int, c_int32_t, c_double_complex
       integer(c_int32_t),value :: ispin, iorb, L
       complex(c_double_complex),intent(in) :: z(*)
       complex(c_double_complex),intent(out) :: delta(*)
     end function dmed_delta
     integer(c_int) function dmed_chi2_fit(scheme,weight,fg,L,beta,xmu,hloc_aa,nbath,e,v,&
          niter,ftol,chi2,iter) bind(C,name="dmed_chi2_fit")
       import :: c_int, c_int32_t, c_double, c_double_complex
       integer(c_int32_t),value :: scheme, weight, L, nbath, niter
       complex(c_double_complex),intent(in) :: fg(*)
       real(c_double),value :: beta, xmu, hloc_aa, ftol
       real(c_double),intent(inout) :: e(*), v(*)
       real(c_double),intent(out) :: chi2
       integer(c_int32_t),intent(out) :: iter
     end function dmed_chi2_fit
     integer(c_int) function dmed_chi2_fit_hybrid(scheme,weight,fg,L,beta,xmu,hloc,norb,nbath,&
          e,v,niter,ftol,chi2,iter) bind(C,name="dmed_chi2_fit_hybrid")
       import :: c_int, c_int32_t, c_double, c_double_complex
       integer(c_int32_t),value :: scheme, weight, L, norb, nbath, niter
       complex(c_double_complex),intent(in) :: fg(*)
       real(c_double),value :: beta, xmu, ftol
       real(c_double),intent(in) :: hloc(*)
       real(c_double),intent(inout) :: e(*), v(*)
       real(c_double),intent(out) :: chi2
       integer(c_int32_t),intent(out) :: iter
     end function dmed_chi2_fit_hybrid
  end interface

contains

  !> copy a NUL-terminated C string into a Fortran character variable
  function c_string(p) result(s)
    type(c_ptr),intent(in) :: p
    character(len=256)     :: s
    character(kind=c_char),pointer :: a(:)
    integer :: i
    s = ""
    if(.not.c_associated(p))return
    call c_f_pointer(p,a,[256])
    do i=1,256
       if(a(i)==c_null_char)exit
       s(i:i)=a(i)
    enddo
  end function c_string

end module DMED_CABI
