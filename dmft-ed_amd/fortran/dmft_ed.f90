!> DMFT_ED -- Fortran API of the impurity solver, same names, arguments and
!> input parameters as the reference module (DMFT_ED.f90:1-75), so that an
!> existing driver relinks unchanged.  The work is done by the gfx950
!> library libdmed.so through ISO_C_BINDING (module DMED_CABI).
!>
!>   ed_read_input            ED_INPUT_VARS.f90:93-202   (same names/defaults)
!>   ed_init_solver           ED_MAIN.f90:67-107
!>   ed_solve                 ED_MAIN.f90:251-289
!>   ed_get_sigma_matsubara   ED_IO/get_sigma_matsubara.f90:2-5
!>   ed_get_sigma_real, ed_get_gimp_matsubara/real, ed_get_dens, ed_get_docc
!>   ed_chi2_fitgf            ED_CHI2FIT.f90:86-138 (normal bath)
!>   get_bath_dimension       ED_BATH/user_aux.f90:11-30
!>
!>   lattice forms (inequivalent sites, ED_MAIN.f90:170-195, 364-468;
!>   ED_CHI2FIT.f90:413-452; ED_IO get_*_lattice): ed_init_solver,
!>   ed_solve, ed_chi2_fitgf and the getters with a leading site index,
!>   the sites solved concurrently on the GPU
!>
!> Scope: ed_mode="normal", bath_type="normal".  Other modes stop with a
!> message, as the reference stops on unsupported input.
module DMFT_ED
  use iso_c_binding
  use DMED_CABI
  implicit none
  private

  !-------------------------------------------------- input variables
  integer,public              :: Norb=1, Nbath=6, Nspin=1, nloop=100
  real(8),public              :: Uloc(3)=[2d0,0d0,0d0], Ust=0d0, Jh=0d0, Jx=0d0, Jp=0d0
  real(8),public              :: beta=1000d0, xmu=0d0, deltasc=0.02d0, dmft_error=1d-5
  real(8),public              :: sb_field=0.1d0, eps=0.01d0, wini=-5d0, wfin=5d0
  integer,public              :: Nsuccess=1, Lmats=5000, Lreal=5000, Ltau=1000, Lfit=1000
  real(8),public              :: nread=0d0, nerr=1d-4, ndelta=0.1d0
  logical,public              :: ed_twin=.false., ed_sparse_H=.true., chiflag=.false.
  logical,public              :: Jhflag=.false., HFmode=.true.
  logical,public              :: ed_print_Sigma=.true., ed_print_G=.true., ed_print_G0=.true.
  real(8),public              :: cutoff=1d-9, gs_threshold=1d-9, hwband=2d0
  integer,public              :: lanc_nstates_sector=6, lanc_nstates_total=1, lanc_nstates_step=2
  integer,public              :: lanc_ncv_factor=3, lanc_ncv_add=5, lanc_niter=512, lanc_ngfiter=200
  real(8),public              :: lanc_tolerance=1d-12
  integer,public              :: lanc_dim_threshold=256
  integer,public              :: cg_niter=500, cg_method=0, cg_stop=0, cg_weight=0
  real(8),public              :: cg_ftol=1d-5, cg_eps=1d-6
  character(len=5),public     :: cg_scheme='weiss'
  character(len=7),public     :: ed_mode='normal', bath_type='normal'
  logical,public              :: ed_para=.true.
  real(8),public              :: ed_vsf_ratio=0.1d0, ed_bath_noise_thr=0d0
  character(len=100),public   :: Hfile='hamiltonian', HLOCfile='inputHLOC.in'
  integer,public              :: LOGfile=6, ed_verbose=3
  !> GPU-path extensions (not in the reference): FP32 Green's-function chains
  !> and the batched sector engine
  logical,public              :: ed_gpu_fp32_gf=.false., ed_gpu_batched=.false.
  logical,public              :: ed_gpu_mixed_gs=.false.

  !-------------------------------------------------- solver state
  complex(8),allocatable,public :: impGmats(:,:,:,:,:), impGreal(:,:,:,:,:)
  complex(8),allocatable,public :: impSmats(:,:,:,:,:), impSreal(:,:,:,:,:)
  real(8),allocatable,public    :: ed_dens(:), ed_docc(:)
  real(8),allocatable,public    :: ed_dens_up(:), ed_dens_dw(:)
  real(8),public                :: ed_egs = 0d0
  !> local energies and double occupancies of the last solve
  !> (local_energy_impurity, ED_OBSERVABLES.f90:450-669)
  real(8),public                :: ed_Epot=0d0, ed_Eint=0d0, ed_Ehartree=0d0, ed_Eknot=0d0
  real(8),public                :: ed_Dust=0d0, ed_Dund=0d0, ed_Dse=0d0, ed_Dph=0d0
  type(dmed_local_obs_t),target :: ed_obs
  !> spin susceptibility of the last solve with CHIFLAG (build_chi_spin.f90):
  !> rows 1..Norb per orbital, Norb+1 the total (Norb > 1)
  real(8),allocatable,target,public    :: spinChi_tau(:,:)
  complex(8),allocatable,target,public :: spinChi_iv(:,:), spinChi_w(:,:)
  !> charge susceptibility of the last solve with CHIFLAG (build_chi_dens.f90):
  !> (Norb,Norb,0:Ltau) / (Norb,Norb,Lreal) / (Norb,Norb,0:Lmats), and the totals
  real(8),allocatable,target,public    :: densChi_tau(:,:,:), densChi_mix_tau(:,:,:)
  complex(8),allocatable,target,public :: densChi_iv(:,:,:), densChi_w(:,:,:)
  complex(8),allocatable,target,public :: densChi_mix_iv(:,:,:), densChi_mix_w(:,:,:)
  real(8),allocatable,target,public    :: densChi_tot_tau(:)
  complex(8),allocatable,target,public :: densChi_tot_iv(:), densChi_tot_w(:)
  real(8),allocatable,target    :: bath_e(:,:,:), bath_v(:,:,:)
  real(8),allocatable,target    :: impHloc(:,:,:,:)
  integer(c_int32_t),allocatable,target :: neigen_sector(:)
  logical                       :: isetup=.true.
  !> file-name suffix of the site being handled (ed_file_suffix of the
  !> reference): "" for a single site, "_ineq<4-digit site>" for the lattice
  character(len=16)             :: ed_file_suffix=""
  !> lattice results (ED_VARS_GLOBAL.f90 Smatsii, nii, ...): site index first
  complex(8),allocatable        :: Smatsii(:,:,:,:,:,:), Srealii(:,:,:,:,:,:)
  complex(8),allocatable        :: Gmatsii(:,:,:,:,:,:), Grealii(:,:,:,:,:,:)
  real(8),allocatable           :: nii(:,:), dii(:,:), mii(:,:), eii(:,:), ddii(:,:)
  integer(c_int32_t),allocatable,target :: neigen_sectorii(:,:), neigen_totalii(:)

  interface ed_init_solver
     module procedure ed_init_solver_single, ed_init_solver_lattice
  end interface ed_init_solver
  interface set_Hloc
     module procedure set_hloc_nn, set_hloc_so
  end interface set_Hloc
  interface lso2nnn_reshape
     module procedure d_nlso2nnn, c_nlso2nnn
  end interface lso2nnn_reshape
  interface nnn2lso_reshape
     module procedure d_nnn2nlso, c_nnn2nlso
  end interface nnn2lso_reshape
  interface so2nn_reshape
     module procedure d_nso2nn, c_nso2nn
  end interface so2nn_reshape
  interface nn2so_reshape
     module procedure d_nn2nso, c_nn2nso
  end interface nn2so_reshape
  interface spin_symmetrize_bath
     module procedure spin_symmetrize_bath_site, spin_symmetrize_bath_lattice
  end interface spin_symmetrize_bath
  interface break_symmetry_bath
     module procedure break_symmetry_bath_site, break_symmetry_bath_lattice
  end interface break_symmetry_bath
  interface ph_symmetrize_bath
     module procedure ph_symmetrize_bath_site, ph_symmetrize_bath_lattice
  end interface ph_symmetrize_bath
  interface ph_trans_bath
     module procedure ph_trans_bath_site, ph_trans_bath_lattice
  end interface ph_trans_bath
  interface enforce_normal_bath
     module procedure enforce_normal_bath_site, enforce_normal_bath_lattice
  end interface enforce_normal_bath
  interface ed_solve
     module procedure ed_solve_single, ed_solve_lattice
  end interface ed_solve
  interface ed_chi2_fitgf
     module procedure ed_chi2_fitgf_single, ed_fit_bath_sites_normal
  end interface ed_chi2_fitgf
  interface ed_get_self_matsubara
     module procedure anomalous_zero_1, anomalous_zero_2, anomalous_zero_3
     module procedure anomalous_zero_lattice_1, anomalous_zero_lattice_2, anomalous_zero_lattice_3
     module procedure anomalous_zero_lattice_11, anomalous_zero_lattice_21
  end interface ed_get_self_matsubara
  interface ed_get_self_real
     module procedure anomalous_zero_1, anomalous_zero_2, anomalous_zero_3
     module procedure anomalous_zero_lattice_1, anomalous_zero_lattice_2, anomalous_zero_lattice_3
     module procedure anomalous_zero_lattice_11, anomalous_zero_lattice_21
  end interface ed_get_self_real
  interface ed_get_fimp_matsubara
     module procedure anomalous_zero_1, anomalous_zero_2, anomalous_zero_3
     module procedure anomalous_zero_lattice_1, anomalous_zero_lattice_2, anomalous_zero_lattice_3
     module procedure anomalous_zero_lattice_11, anomalous_zero_lattice_21
  end interface ed_get_fimp_matsubara
  interface ed_get_fimp_real
     module procedure anomalous_zero_1, anomalous_zero_2, anomalous_zero_3
     module procedure anomalous_zero_lattice_1, anomalous_zero_lattice_2, anomalous_zero_lattice_3
     module procedure anomalous_zero_lattice_11, anomalous_zero_lattice_21
  end interface ed_get_fimp_real
  interface ed_get_phisc
     module procedure ed_get_phisc_1, ed_get_phisc_2, ed_get_phisc_lattice_1, ed_get_phisc_lattice_2
  end interface ed_get_phisc
  interface get_component_bath
     module procedure get_full_component_bath, get_spin_component_bath, get_spin_orb_component_bath
  end interface get_component_bath
  interface set_component_bath
     module procedure set_full_component_bath, set_spin_component_bath, set_spin_orb_component_bath
  end interface set_component_bath
  interface copy_component_bath
     module procedure copy_full_component_bath, copy_spin_component_bath, copy_spin_orb_component_bath
  end interface copy_component_bath
  interface ed_get_sigma_matsubara
     module procedure ed_get_sigma_matsubara_1, ed_get_sigma_matsubara_2, ed_get_sigma_matsubara_3
     module procedure ed_get_sigma_matsubara_lattice_1, ed_get_sigma_matsubara_lattice_2, ed_get_sigma_matsubara_lattice_3
     module procedure ed_get_sigma_matsubara_lattice_11, ed_get_sigma_matsubara_lattice_21, ed_get_sigma_matsubara_lattice_31
  end interface ed_get_sigma_matsubara
  interface ed_get_sigma_real
     module procedure ed_get_sigma_real_1, ed_get_sigma_real_2, ed_get_sigma_real_3
     module procedure ed_get_sigma_real_lattice_1, ed_get_sigma_real_lattice_2, ed_get_sigma_real_lattice_3
     module procedure ed_get_sigma_real_lattice_11, ed_get_sigma_real_lattice_21, ed_get_sigma_real_lattice_31
  end interface ed_get_sigma_real
  interface ed_get_gimp_matsubara
     module procedure ed_get_gimp_matsubara_1, ed_get_gimp_matsubara_2, ed_get_gimp_matsubara_3
     module procedure ed_get_gimp_matsubara_lattice_1, ed_get_gimp_matsubara_lattice_2, ed_get_gimp_matsubara_lattice_3
     module procedure ed_get_gimp_matsubara_lattice_11, ed_get_gimp_matsubara_lattice_21, ed_get_gimp_matsubara_lattice_31
  end interface ed_get_gimp_matsubara
  interface ed_get_gimp_real
     module procedure ed_get_gimp_real_1, ed_get_gimp_real_2, ed_get_gimp_real_3
     module procedure ed_get_gimp_real_lattice_1, ed_get_gimp_real_lattice_2, ed_get_gimp_real_lattice_3
     module procedure ed_get_gimp_real_lattice_11, ed_get_gimp_real_lattice_21, ed_get_gimp_real_lattice_31
  end interface ed_get_gimp_real
  interface ed_get_dens
     module procedure ed_get_dens_1, ed_get_dens_2, ed_get_dens_lattice_1, ed_get_dens_lattice_2
  end interface ed_get_dens
  interface ed_get_docc
     module procedure ed_get_docc_1, ed_get_docc_2, ed_get_docc_lattice_1, ed_get_docc_lattice_2
  end interface ed_get_docc
  interface ed_get_mag
     module procedure ed_get_mag_1, ed_get_mag_2, ed_get_mag_lattice_1, ed_get_mag_lattice_2
  end interface ed_get_mag
  interface ed_get_eimp
     module procedure ed_get_eimp_1, ed_get_eimp_lattice
  end interface ed_get_eimp
  interface ed_get_doubles
     module procedure ed_get_doubles_1, ed_get_doubles_lattice
  end interface ed_get_doubles

  public :: ed_read_input
  public :: get_bath_dimension
  public :: ed_init_solver
  public :: ed_push_model, set_bath, ed_check
  public :: ed_solve
  public :: ed_get_sigma_matsubara, ed_get_sigma_real
  public :: ed_get_gimp_matsubara, ed_get_gimp_real
  public :: ed_get_dens, ed_get_docc
  public :: ed_get_dens_up, ed_get_dens_dw, ed_get_mag
  public :: ed_get_eimp, ed_get_epot, ed_get_eint, ed_get_ehartree, ed_get_eknot
  public :: ed_get_doubles, ed_get_dust, ed_get_dund, ed_get_dse, ed_get_dph
  public :: ed_chi2_fitgf
  public :: ed_print_impSigma, ed_print_impG, ed_print_impG0, ed_print_impChi
  public :: ed_read_impSigma
  public :: set_Hloc, lso2nnn_reshape, nnn2lso_reshape, so2nn_reshape, nn2so_reshape
  public :: spin_symmetrize_bath, break_symmetry_bath, ph_symmetrize_bath, ph_trans_bath
  public :: enforce_normal_bath
  public :: ed_get_self_matsubara, ed_get_self_real, ed_get_fimp_matsubara, ed_get_fimp_real
  public :: ed_get_phisc
  public :: get_component_bath_dimension, get_spin_component_bath_dimension
  public :: get_orb_component_bath_dimension, get_spin_orb_component_bath_dimension
  public :: get_component_bath, set_component_bath, copy_component_bath
  public :: so2os_reshape, os2so_reshape

contains

  !---------------------------------------------------------------- input
  !> Read NAME=value lines (comments after '!').  Names and defaults are
  !> those of ED_INPUT_VARS.f90:111-176; unknown names are ignored.
  subroutine ed_read_input(INPUTunit,comm)
    character(len=*)  :: INPUTunit
    integer,optional  :: comm
    integer           :: unit, ios, ieq, iex
    character(len=512):: line, key, val
    logical           :: exists
    inquire(file=trim(INPUTunit),exist=exists)
    if(.not.exists)then
       write(LOGfile,"(A)")"ed_read_input: "//trim(INPUTunit)//" not found, using defaults"
       return
    endif
    open(newunit=unit,file=trim(INPUTunit),status='old',action='read')
    do
       read(unit,"(A)",iostat=ios)line
       if(ios/=0)exit
       iex = index(line,'!')
       if(iex>0)line=line(:iex-1)
       ieq = index(line,'=')
       if(ieq<=1)cycle
       key = upper(adjustl(line(:ieq-1)))
       val = adjustl(line(ieq+1:))
       call set_var(trim(key),trim(val))
    enddo
    close(unit)
    Ltau=max(int(beta),Ltau)
    if(Lfit>Lmats)Lfit=Lmats
  end subroutine ed_read_input

  subroutine set_var(key,val)
    character(len=*) :: key, val
    integer          :: ios
    character(len=len(val)+2) :: v
    v = val
    select case(key)
    case("NORB");  read(v,*,iostat=ios)Norb
    case("NBATH"); read(v,*,iostat=ios)Nbath
    case("NSPIN"); read(v,*,iostat=ios)Nspin
    case("ULOC");  read(v,*,iostat=ios)Uloc
    case("UST");   read(v,*,iostat=ios)Ust
    case("JH");    read(v,*,iostat=ios)Jh
    case("JX");    read(v,*,iostat=ios)Jx
    case("JP");    read(v,*,iostat=ios)Jp
    case("BETA");  read(v,*,iostat=ios)beta
    case("XMU");   read(v,*,iostat=ios)xmu
    case("DELTASC"); read(v,*,iostat=ios)deltasc
    case("NLOOP"); read(v,*,iostat=ios)nloop
    case("DMFT_ERROR"); read(v,*,iostat=ios)dmft_error
    case("SB_FIELD"); read(v,*,iostat=ios)sb_field
    case("ED_TWIN"); read(v,*,iostat=ios)ed_twin
    case("ED_SPARSE_H"); read(v,*,iostat=ios)ed_sparse_H
    case("ED_PRINT_SIGMA"); read(v,*,iostat=ios)ed_print_Sigma
    case("ED_PRINT_G"); read(v,*,iostat=ios)ed_print_G
    case("ED_PRINT_G0"); read(v,*,iostat=ios)ed_print_G0
    case("NSUCCESS"); read(v,*,iostat=ios)Nsuccess
    case("LMATS"); read(v,*,iostat=ios)Lmats
    case("LREAL"); read(v,*,iostat=ios)Lreal
    case("LTAU");  read(v,*,iostat=ios)Ltau
    case("LFIT");  read(v,*,iostat=ios)Lfit
    case("NREAD"); read(v,*,iostat=ios)nread
    case("NERR");  read(v,*,iostat=ios)nerr
    case("NDELTA"); read(v,*,iostat=ios)ndelta
    case("WINI");  read(v,*,iostat=ios)wini
    case("WFIN");  read(v,*,iostat=ios)wfin
    case("CHIFLAG"); read(v,*,iostat=ios)chiflag
    case("JHFLAG"); read(v,*,iostat=ios)Jhflag
    case("HFMODE"); read(v,*,iostat=ios)HFmode
    case("EPS");   read(v,*,iostat=ios)eps
    case("CUTOFF"); read(v,*,iostat=ios)cutoff
    case("GS_THRESHOLD"); read(v,*,iostat=ios)gs_threshold
    case("HWBAND"); read(v,*,iostat=ios)hwband
    case("LANC_NSTATES_SECTOR"); read(v,*,iostat=ios)lanc_nstates_sector
    case("LANC_NSTATES_TOTAL"); read(v,*,iostat=ios)lanc_nstates_total
    case("LANC_NSTATES_STEP"); read(v,*,iostat=ios)lanc_nstates_step
    case("LANC_NCV_FACTOR"); read(v,*,iostat=ios)lanc_ncv_factor
    case("LANC_NCV_ADD"); read(v,*,iostat=ios)lanc_ncv_add
    case("LANC_NITER"); read(v,*,iostat=ios)lanc_niter
    case("LANC_NGFITER"); read(v,*,iostat=ios)lanc_ngfiter
    case("LANC_TOLERANCE"); read(v,*,iostat=ios)lanc_tolerance
    case("LANC_DIM_THRESHOLD"); read(v,*,iostat=ios)lanc_dim_threshold
    case("CG_NITER"); read(v,*,iostat=ios)cg_niter
    case("CG_SCHEME"); cg_scheme=trim(adjustl(v))
    case("CG_FTOL"); read(v,*,iostat=ios)cg_ftol
    case("CG_METHOD"); read(v,*,iostat=ios)cg_method
    case("CG_STOP"); read(v,*,iostat=ios)cg_stop
    case("CG_EPS"); read(v,*,iostat=ios)cg_eps
    case("CG_WEIGHT"); read(v,*,iostat=ios)cg_weight
    case("ED_MODE"); ed_mode=trim(adjustl(v))
    case("ED_PARA"); read(v,*,iostat=ios)ed_para
    case("ED_VSF_RATIO"); read(v,*,iostat=ios)ed_vsf_ratio
    case("ED_BATH_NOISE_THR"); read(v,*,iostat=ios)ed_bath_noise_thr
    case("BATH_TYPE"); bath_type=trim(adjustl(v))
    case("HFILE"); Hfile=trim(adjustl(v))
    case("IMPHFILE"); HLOCfile=trim(adjustl(v))
    case("LOGFILE"); read(v,*,iostat=ios)LOGfile
    case("ED_VERBOSE"); read(v,*,iostat=ios)ed_verbose
    case("ED_GPU_FP32_GF"); read(v,*,iostat=ios)ed_gpu_fp32_gf
    case("ED_GPU_BATCHED"); read(v,*,iostat=ios)ed_gpu_batched
    case("ED_GPU_MIXED_GS"); read(v,*,iostat=ios)ed_gpu_mixed_gs
    case default
    end select
  end subroutine set_var

  function upper(s) result(u)
    character(len=*) :: s
    character(len=len(s)) :: u
    integer :: i, c
    u = s
    do i=1,len(s)
       c = iachar(s(i:i))
       if(c>=97 .and. c<=122)u(i:i)=achar(c-32)
    enddo
  end function upper

  !------------------------------------------------------------- bath
  !> ( e [Nspin][Norb][Nbath] + v [Nspin][Norb][Nbath] ), user_aux.f90:17-30
  function get_bath_dimension(Hloc_nn,ispin_) result(bath_size)
    complex(8),optional,intent(in) :: Hloc_nn(:,:,:,:)
    integer,optional               :: ispin_
    integer                        :: bath_size
    if(ed_mode/='normal')stop "get_bath_dimension: only ed_mode=normal"
    select case(trim(bath_type))
    case('normal')
       bath_size = Norb*Nbath + Norb*Nbath
    case('hybrid')
       !(e [Nspin][1][Nbath] + v [Nspin][Norb][Nbath]) (ED_BATH/user_aux.f90:31-35)
       bath_size = Nbath + Norb*Nbath
    case default
       stop "get_bath_dimension: only bath_type=normal or hybrid"
    end select
    if(.not.present(ispin_))bath_size=Nspin*bath_size
  end function get_bath_dimension

  subroutine check_modes()
    if(ed_mode/='normal')stop "DMFT_ED: only ed_mode=normal is built on the GPU path"
    if(bath_type/='normal' .and. bath_type/='hybrid')stop "DMFT_ED: only bath_type=normal or hybrid is built on the GPU path"
    if(Nspin>2)stop "ED ERROR: Nspin > 2 is currently not supported"
    if(Norb>3)stop "ED ERROR: Norb > 3 is currently not supported"
  end subroutine check_modes

  !> set_dmft_bath (dmft_aux.f90:489-512; hybrid :572-590): flat bath ->
  !> e(ispin,iorb,k), v(...); the hybrid bath keeps its levels in e(ispin,1,k)
  subroutine set_bath(bath)
    real(8),intent(in) :: bath(:)
    integer :: s, a, k, io, st
    if(trim(bath_type)=='hybrid')then
       do s=1,Nspin
          do k=1,Nbath
             bath_e(s,:,k) = bath(k + (s-1)*Nbath)
          enddo
          do a=1,Norb
             do k=1,Nbath
                bath_v(s,a,k) = bath(Nspin*Nbath + k + (a-1)*Nbath + (s-1)*Norb*Nbath)
             enddo
          enddo
       enddo
       return
    endif
    st = Nspin*Norb*Nbath
    do s=1,Nspin
       do a=1,Norb
          do k=1,Nbath
             io = k + (a-1)*Nbath + (s-1)*Nbath*Norb
             bath_e(s,a,k) = bath(io)
             bath_v(s,a,k) = bath(st+io)
          enddo
       enddo
    enddo
  end subroutine set_bath

  subroutine get_bath(bath)
    real(8),intent(out) :: bath(:)
    integer :: s, a, k, io, st
    if(trim(bath_type)=='hybrid')then
       do s=1,Nspin
          do k=1,Nbath
             bath(k + (s-1)*Nbath) = bath_e(s,1,k)
          enddo
          do a=1,Norb
             do k=1,Nbath
                bath(Nspin*Nbath + k + (a-1)*Nbath + (s-1)*Norb*Nbath) = bath_v(s,a,k)
             enddo
          enddo
       enddo
       return
    endif
    st = Nspin*Norb*Nbath
    do s=1,Nspin
       do a=1,Norb
          do k=1,Nbath
             io = k + (a-1)*Nbath + (s-1)*Nbath*Norb
             bath(io)    = bath_e(s,a,k)
             bath(st+io) = bath_v(s,a,k)
          enddo
       enddo
    enddo
  end subroutine get_bath

  !> init_dmft_bath, normal bath (dmft_aux.f90:92-177): levels spread over
  !> [-hwband, hwband], V = max(0.1, 1/sqrt(Nbath)); then the
  !> hamiltonian.restart file if present.
  subroutine init_bath()
    integer :: i, Nh, unit, s, a, ios
    real(8) :: de
    logical :: exists
    character(len=512) :: line
    bath_e(:,:,1)     = -hwband
    bath_e(:,:,Nbath) =  hwband
    Nh = Nbath/2
    if(mod(Nbath,2)==0 .and. Nbath>=4)then
       de = hwband/max(Nh-1,1)
       bath_e(:,:,Nh)   = -1d-3
       bath_e(:,:,Nh+1) =  1d-3
       do i=2,Nh-1
          bath_e(:,:,i)         = -hwband + (i-1)*de
          bath_e(:,:,Nbath-i+1) =  hwband - (i-1)*de
       enddo
    elseif(mod(Nbath,2)/=0 .and. Nbath>=3)then
       de = hwband/Nh
       bath_e(:,:,Nh+1) = 0d0
       do i=2,Nh
          bath_e(:,:,i)         = -hwband + (i-1)*de
          bath_e(:,:,Nbath-i+1) =  hwband - (i-1)*de
       enddo
    endif
    do i=1,Nbath
       bath_v(:,:,i) = max(0.1d0,1d0/sqrt(dble(Nbath)))
    enddo
    inquire(file=trim(Hfile)//trim(ed_file_suffix)//".restart",exist=exists)
    if(exists)then
       write(LOGfile,"(A)")"Reading bath from file "//trim(Hfile)//trim(ed_file_suffix)//".restart"
       open(newunit=unit,file=trim(Hfile)//trim(ed_file_suffix)//".restart",status='old',action='read')
       read(unit,"(A)",iostat=ios)line
       do i=1,Nbath
          if(trim(bath_type)=='hybrid')then
             ! (e(s,1,i), (v(s,a,i), a=1,Norb), s=1,Nspin) (dmft_aux.f90:196-205)
             read(unit,*,iostat=ios)(bath_e(s,1,i),(bath_v(s,a,i),a=1,Norb),s=1,Nspin)
             do s=1,Nspin
                bath_e(s,:,i) = bath_e(s,1,i)
             enddo
          else
             read(unit,*,iostat=ios)((bath_e(s,a,i),bath_v(s,a,i),a=1,Norb),s=1,Nspin)
          endif
          if(ios/=0)exit
       enddo
       close(unit)
    endif
  end subroutine init_bath

  !> save_dmft_bath (dmft_aux.f90:446-465): hamiltonian.used / .restart
  subroutine save_bath(used)
    logical,intent(in) :: used
    integer :: unit, i, s, a
    character(len=16) :: ext
    ext = ".restart"; if(used)ext = ".used"
    open(newunit=unit,file=trim(Hfile)//trim(ed_file_suffix)//trim(ext))
    if(trim(bath_type)=='hybrid')then
       ! write_dmft_bath, hybrid (dmft_aux.f90:366-377)
       write(unit,"(90(A21,1X))")("#Ek_s"//itoa(s),("Vk_l"//itoa(a)//"_s"//itoa(s),a=1,Norb),s=1,Nspin)
       do i=1,Nbath
          write(unit,"(90(F21.12,1X))")(bath_e(s,1,i),(bath_v(s,a,i),a=1,Norb),s=1,Nspin)
       enddo
    else
       write(unit,"(90(A21,1X))")(("#Ek_l"//itoa(a)//"_s"//itoa(s),"Vk_l"//itoa(a)//"_s"//itoa(s),a=1,Norb),s=1,Nspin)
       do i=1,Nbath
          write(unit,"(90(F21.12,1X))")((bath_e(s,a,i),bath_v(s,a,i),a=1,Norb),s=1,Nspin)
       enddo
    endif
    close(unit)
  end subroutine save_bath

  function itoa(i) result(s)
    integer :: i
    character(len=:),allocatable :: s
    character(len=16) :: b
    write(b,"(I0)")i
    s = trim(b)
  end function itoa

  !------------------------------------------------------------ solver
  subroutine ed_check(rc,where)
    integer(c_int),intent(in) :: rc
    character(len=*) :: where
    if(rc/=0)then
       write(*,"(A)")"DMFT_ED "//where//" error: "//trim(c_string(dmed_last_error()))
       error stop "DMFT_ED: GPU library error"
    endif
  end subroutine ed_check

  !> set_Hloc_nn / set_Hloc_so (ED_AUX_FUNX.f90:135-153)
  subroutine set_hloc_nn(Hloc)
    complex(8),intent(in) :: Hloc(Nspin,Nspin,Norb,Norb)
    if(any(abs(aimag(Hloc))>0d0))stop "DMFT_ED: complex Hloc is not supported on the GPU path (H must be real)"
    if(.not.allocated(impHloc))allocate(impHloc(Nspin,Nspin,Norb,Norb))
    impHloc = dble(Hloc)
  end subroutine set_hloc_nn

  subroutine set_hloc_so(Hloc)
    complex(8),intent(in) :: Hloc(Nspin*Norb,Nspin*Norb)
    call set_hloc_nn(so2nn_reshape(Hloc,Nspin,Norb))
  end subroutine set_hloc_so

  !-------------------------------------------- reshapes (ED_AUX_FUNX.f90:177-351)
  !> (Nlat*Nspin*Norb)^2 <-> (Nlat,Nspin,Nspin,Norb,Norb), index
  !> io = iorb + (ispin-1)*Norb + (ilat-1)*Norb*Nspin, site-diagonal blocks
  function d_nlso2nnn(Hlso,Nlat,Nspin_,Norb_) result(Hnnn)
    integer :: Nlat, Nspin_, Norb_
    real(8),dimension(Nlat*Nspin_*Norb_,Nlat*Nspin_*Norb_) :: Hlso
    real(8),dimension(Nlat,Nspin_,Nspin_,Norb_,Norb_)      :: Hnnn
    integer :: l, s1, s2, a, b
    do l=1,Nlat
       do s1=1,Nspin_
          do s2=1,Nspin_
             do a=1,Norb_
                do b=1,Norb_
                   Hnnn(l,s1,s2,a,b) = Hlso(a+(s1-1)*Norb_+(l-1)*Norb_*Nspin_, b+(s2-1)*Norb_+(l-1)*Norb_*Nspin_)
                enddo
             enddo
          enddo
       enddo
    enddo
  end function d_nlso2nnn

  function c_nlso2nnn(Hlso,Nlat,Nspin_,Norb_) result(Hnnn)
    integer :: Nlat, Nspin_, Norb_
    complex(8),dimension(Nlat*Nspin_*Norb_,Nlat*Nspin_*Norb_) :: Hlso
    complex(8),dimension(Nlat,Nspin_,Nspin_,Norb_,Norb_)      :: Hnnn
    integer :: l, s1, s2, a, b
    do l=1,Nlat
       do s1=1,Nspin_
          do s2=1,Nspin_
             do a=1,Norb_
                do b=1,Norb_
                   Hnnn(l,s1,s2,a,b) = Hlso(a+(s1-1)*Norb_+(l-1)*Norb_*Nspin_, b+(s2-1)*Norb_+(l-1)*Norb_*Nspin_)
                enddo
             enddo
          enddo
       enddo
    enddo
  end function c_nlso2nnn

  function d_nnn2nlso(Hnnn,Nlat,Nspin_,Norb_) result(Hlso)
    integer :: Nlat, Nspin_, Norb_
    real(8),dimension(Nlat,Nspin_,Nspin_,Norb_,Norb_)      :: Hnnn
    real(8),dimension(Nlat*Nspin_*Norb_,Nlat*Nspin_*Norb_) :: Hlso
    integer :: l, s1, s2, a, b
    Hlso = 0d0
    do l=1,Nlat
       do s1=1,Nspin_
          do s2=1,Nspin_
             do a=1,Norb_
                do b=1,Norb_
                   Hlso(a+(s1-1)*Norb_+(l-1)*Norb_*Nspin_, b+(s2-1)*Norb_+(l-1)*Norb_*Nspin_) = Hnnn(l,s1,s2,a,b)
                enddo
             enddo
          enddo
       enddo
    enddo
  end function d_nnn2nlso

  function c_nnn2nlso(Hnnn,Nlat,Nspin_,Norb_) result(Hlso)
    integer :: Nlat, Nspin_, Norb_
    complex(8),dimension(Nlat,Nspin_,Nspin_,Norb_,Norb_)      :: Hnnn
    complex(8),dimension(Nlat*Nspin_*Norb_,Nlat*Nspin_*Norb_) :: Hlso
    integer :: l, s1, s2, a, b
    Hlso = (0d0,0d0)
    do l=1,Nlat
       do s1=1,Nspin_
          do s2=1,Nspin_
             do a=1,Norb_
                do b=1,Norb_
                   Hlso(a+(s1-1)*Norb_+(l-1)*Norb_*Nspin_, b+(s2-1)*Norb_+(l-1)*Norb_*Nspin_) = Hnnn(l,s1,s2,a,b)
                enddo
             enddo
          enddo
       enddo
    enddo
  end function c_nnn2nlso

  function d_nso2nn(Hso,Nspin_,Norb_) result(Hnn)
    integer :: Nspin_, Norb_
    real(8),dimension(Nspin_*Norb_,Nspin_*Norb_) :: Hso
    real(8),dimension(Nspin_,Nspin_,Norb_,Norb_) :: Hnn
    real(8) :: t(1,Nspin_,Nspin_,Norb_,Norb_)
    t = d_nlso2nnn(Hso,1,Nspin_,Norb_)
    Hnn = t(1,:,:,:,:)
  end function d_nso2nn

  function c_nso2nn(Hso,Nspin_,Norb_) result(Hnn)
    integer :: Nspin_, Norb_
    complex(8),dimension(Nspin_*Norb_,Nspin_*Norb_) :: Hso
    complex(8),dimension(Nspin_,Nspin_,Norb_,Norb_) :: Hnn
    complex(8) :: t(1,Nspin_,Nspin_,Norb_,Norb_)
    t = c_nlso2nnn(Hso,1,Nspin_,Norb_)
    Hnn = t(1,:,:,:,:)
  end function c_nso2nn

  function d_nn2nso(Hnn,Nspin_,Norb_) result(Hso)
    integer :: Nspin_, Norb_
    real(8),dimension(Nspin_,Nspin_,Norb_,Norb_) :: Hnn
    real(8),dimension(Nspin_*Norb_,Nspin_*Norb_) :: Hso
    real(8) :: t(1,Nspin_,Nspin_,Norb_,Norb_)
    t(1,:,:,:,:) = Hnn
    Hso = d_nnn2nlso(t,1,Nspin_,Norb_)
  end function d_nn2nso

  function c_nn2nso(Hnn,Nspin_,Norb_) result(Hso)
    integer :: Nspin_, Norb_
    complex(8),dimension(Nspin_,Nspin_,Norb_,Norb_) :: Hnn
    complex(8),dimension(Nspin_*Norb_,Nspin_*Norb_) :: Hso
    complex(8) :: t(1,Nspin_,Nspin_,Norb_,Norb_)
    t(1,:,:,:,:) = Hnn
    Hso = c_nnn2nlso(t,1,Nspin_,Norb_)
  end function c_nn2nso

  !----------------------------------- bath symmetries (ED_BATH/user_aux.f90:1071-1270)
  subroutine bath_arrays()
    if(.not.allocated(bath_e))allocate(bath_e(Nspin,Norb,Nbath), bath_v(Nspin,Norb,Nbath))
  end subroutine bath_arrays

  !> spin-down bath := spin-up bath (Nspin = 2)
  subroutine spin_symmetrize_bath_site(bath,save)
    real(8),dimension(:) :: bath
    logical,optional     :: save
    if(Nspin==1)then
       write(LOGfile,"(A)")"spin_symmetrize_bath: Nspin=1 nothing to symmetrize"
       return
    endif
    call bath_arrays(); call set_bath(bath)
    bath_e(Nspin,:,:) = bath_e(1,:,:)
    bath_v(Nspin,:,:) = bath_v(1,:,:)
    call finish_bath(bath,save)
  end subroutine spin_symmetrize_bath_site

  !> e_up += sign*field, e_dw -= sign*field
  subroutine break_symmetry_bath_site(bath,field,sign,save)
    real(8),dimension(:) :: bath
    real(8)              :: field, sign
    logical,optional     :: save
    call bath_arrays(); call set_bath(bath)
    bath_e(1,:,:) = bath_e(1,:,:) + sign*field
    bath_e(Nspin,:,:) = bath_e(Nspin,:,:) - sign*field
    call finish_bath(bath,save)
  end subroutine break_symmetry_bath_site

  !> e(Nbath+1-i) = -e(i), V(Nbath+1-i) = V(i); the middle level of an odd bath at 0
  subroutine ph_symmetrize_bath_site(bath,save)
    real(8),dimension(:) :: bath
    logical,optional     :: save
    integer :: i
    if(Nbath==1)return
    call bath_arrays(); call set_bath(bath)
    do i=1,Nbath/2
       bath_e(:,:,Nbath+1-i) = -bath_e(:,:,i)
       bath_v(:,:,Nbath+1-i) =  bath_v(:,:,i)
    enddo
    if(mod(Nbath,2)/=0)bath_e(:,:,(Nbath-1)/2+1) = 0d0
    call finish_bath(bath,save)
  end subroutine ph_symmetrize_bath_site

  !> particle-hole transformation of the normal bath: Norb = 1: e -> -e;
  !> Norb = 2: e_1 <-> -e_2, V_1 <-> V_2; other Norb unchanged
  subroutine ph_trans_bath_site(bath,save)
    real(8),dimension(:) :: bath
    logical,optional     :: save
    real(8),allocatable  :: te(:,:), tv(:,:)
    if(Nbath==1)return
    if(trim(bath_type)/='normal')stop "ph_trans_bath: normal bath only"
    call bath_arrays(); call set_bath(bath)
    select case(Norb)
    case(1)
       bath_e = -bath_e
    case(2)
       allocate(te(Nspin,Nbath), tv(Nspin,Nbath))
       te = bath_e(:,2,:); tv = bath_v(:,2,:)
       bath_e(:,2,:) = -bath_e(:,1,:); bath_v(:,2,:) = bath_v(:,1,:)
       bath_e(:,1,:) = -te; bath_v(:,1,:) = tv
    end select
    call finish_bath(bath,save)
  end subroutine ph_trans_bath_site

  !> no pairing amplitudes in the normal mode: the bath is saved unchanged
  subroutine enforce_normal_bath_site(bath,save)
    real(8),dimension(:) :: bath
    logical,optional     :: save
    call bath_arrays(); call set_bath(bath)
    call finish_bath(bath,save)
  end subroutine enforce_normal_bath_site

  subroutine finish_bath(bath,save)
    real(8),dimension(:) :: bath
    logical,optional     :: save
    logical :: save_
    save_ = .true.; if(present(save))save_ = save
    if(save_)call save_bath(used=.false.)
    call get_bath(bath)
  end subroutine finish_bath

  !> the lattice forms loop over sites with the suffix "_site<4-digit site>"
  subroutine spin_symmetrize_bath_lattice(bath,save)
    real(8),dimension(:,:) :: bath
    logical,optional       :: save
    integer :: l
    do l=1,size(bath,1)
       write(ed_file_suffix,"(A,I4.4)")"_site",l
       call spin_symmetrize_bath_site(bath(l,:),save)
    enddo
    ed_file_suffix = ""
  end subroutine spin_symmetrize_bath_lattice

  subroutine break_symmetry_bath_lattice(bath,field,sign,save)
    real(8),dimension(:,:) :: bath
    real(8)                :: field, sign
    logical,optional       :: save
    integer :: l
    do l=1,size(bath,1)
       write(ed_file_suffix,"(A,I4.4)")"_site",l
       call break_symmetry_bath_site(bath(l,:),field,sign,save)
    enddo
    ed_file_suffix = ""
  end subroutine break_symmetry_bath_lattice

  subroutine ph_symmetrize_bath_lattice(bath,save)
    real(8),dimension(:,:) :: bath
    logical,optional       :: save
    integer :: l
    do l=1,size(bath,1)
       write(ed_file_suffix,"(A,I4.4)")"_site",l
       call ph_symmetrize_bath_site(bath(l,:),save)
    enddo
    ed_file_suffix = ""
  end subroutine ph_symmetrize_bath_lattice

  subroutine ph_trans_bath_lattice(bath,save)
    real(8),dimension(:,:) :: bath
    logical,optional       :: save
    integer :: l
    do l=1,size(bath,1)
       write(ed_file_suffix,"(A,I4.4)")"_site",l
       call ph_trans_bath_site(bath(l,:),save)
    enddo
    ed_file_suffix = ""
  end subroutine ph_trans_bath_lattice

  subroutine enforce_normal_bath_lattice(bath,save)
    real(8),dimension(:,:) :: bath
    logical,optional       :: save
    integer :: l
    do l=1,size(bath,1)
       write(ed_file_suffix,"(A,I4.4)")"_site",l
       call enforce_normal_bath_site(bath(l,:),save)
    enddo
    ed_file_suffix = ""
  end subroutine enforce_normal_bath_lattice

  !> The anomalous getters ed_get_self_matsubara/real and
  !> ed_get_fimp_matsubara/real (ED_IO/get_self_matsubara.f90,
  !> get_self_realaxis.f90, get_fimp_matsubara.f90, get_fimp_realaxis.f90):
  !> the anomalous self-energy and Green's function, identically zero in the
  !> normal mode. One set of forms serves all four names: the full
  !> (Nspin,Nspin,Norb,Norb,L) array, the (Nspin*Norb,Nspin*Norb,L) matrix,
  !> one element (L, ispin, jspin, iorb, jorb), and the same three with a
  !> leading site index (Nlat).
  subroutine anomalous_zero_1(F)
    complex(8),intent(inout) :: F(:,:,:,:,:)
    F = (0d0,0d0)
  end subroutine anomalous_zero_1

  subroutine anomalous_zero_2(F)
    complex(8),intent(inout) :: F(:,:,:)
    F = (0d0,0d0)
  end subroutine anomalous_zero_2

  subroutine anomalous_zero_3(F,ispin,jspin,iorb,jorb)
    complex(8),intent(inout) :: F(:)
    integer :: ispin, jspin, iorb, jorb
    if(ispin>Nspin.or.jspin>Nspin.or.iorb>Norb.or.jorb>Norb)stop "ed_get error: index > N_spin or N_orbital"
    F = (0d0,0d0)
  end subroutine anomalous_zero_3

  subroutine anomalous_zero_lattice_1(F,Nsites)
    complex(8),intent(inout) :: F(:,:,:,:,:,:)
    integer :: Nsites
    F(1:Nsites,:,:,:,:,:) = (0d0,0d0)
  end subroutine anomalous_zero_lattice_1

  subroutine anomalous_zero_lattice_2(F,Nsites)
    complex(8),intent(inout) :: F(:,:,:,:)
    integer :: Nsites
    F(1:Nsites,:,:,:) = (0d0,0d0)
  end subroutine anomalous_zero_lattice_2

  subroutine anomalous_zero_lattice_3(F,Nsites,ispin,jspin,iorb,jorb)
    complex(8),intent(inout) :: F(:,:)
    integer :: Nsites, ispin, jspin, iorb, jorb
    if(ispin>Nspin.or.jspin>Nspin.or.iorb>Norb.or.jorb>Norb)stop "ed_get error: index > N_spin or N_orbital"
    F(1:Nsites,:) = (0d0,0d0)
  end subroutine anomalous_zero_lattice_3

  !> One site of the last lattice solve (ED_IO/get_sigma_matsubara.f90:72-100
  !> and the same forms of get_sigma_realaxis, get_gimp_matsubara and
  !> get_gimp_realaxis): _11 the full array, _21 the (Nspin*Norb)^2 matrix
  !> with io = a + (s-1)*Norb, _31 one element.
  subroutine ed_get_sigma_matsubara_lattice_11(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: F
    if(.not.allocated(Smatsii))stop "ed_get_sigma_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Smatsii,1))stop "ed_get_sigma_matsubara error: site index out of range"
    F = Smatsii(ilat,:,:,:,:,:)
  end subroutine ed_get_sigma_matsubara_lattice_11

  subroutine ed_get_sigma_matsubara_lattice_21(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: s1, s2, a, b
    if(.not.allocated(Smatsii))stop "ed_get_sigma_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Smatsii,1))stop "ed_get_sigma_matsubara error: site index out of range"
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb, b+(s2-1)*Norb, :) = Smatsii(ilat,s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_matsubara_lattice_21

  subroutine ed_get_sigma_matsubara_lattice_31(F,ilat,ispin,jspin,iorb,jorb)
    integer :: ilat, ispin, jspin, iorb, jorb
    complex(8),dimension(Lmats),intent(inout) :: F
    if(.not.allocated(Smatsii))stop "ed_get_sigma_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Smatsii,1))stop "ed_get_sigma_matsubara error: site index out of range"
    F = Smatsii(ilat,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_matsubara_lattice_31

  subroutine ed_get_sigma_real_lattice_11(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: F
    if(.not.allocated(Srealii))stop "ed_get_sigma_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Srealii,1))stop "ed_get_sigma_real error: site index out of range"
    F = Srealii(ilat,:,:,:,:,:)
  end subroutine ed_get_sigma_real_lattice_11

  subroutine ed_get_sigma_real_lattice_21(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: s1, s2, a, b
    if(.not.allocated(Srealii))stop "ed_get_sigma_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Srealii,1))stop "ed_get_sigma_real error: site index out of range"
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb, b+(s2-1)*Norb, :) = Srealii(ilat,s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_real_lattice_21

  subroutine ed_get_sigma_real_lattice_31(F,ilat,ispin,jspin,iorb,jorb)
    integer :: ilat, ispin, jspin, iorb, jorb
    complex(8),dimension(Lreal),intent(inout) :: F
    if(.not.allocated(Srealii))stop "ed_get_sigma_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Srealii,1))stop "ed_get_sigma_real error: site index out of range"
    F = Srealii(ilat,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_real_lattice_31

  subroutine ed_get_gimp_matsubara_lattice_11(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: F
    if(.not.allocated(Gmatsii))stop "ed_get_gimp_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Gmatsii,1))stop "ed_get_gimp_matsubara error: site index out of range"
    F = Gmatsii(ilat,:,:,:,:,:)
  end subroutine ed_get_gimp_matsubara_lattice_11

  subroutine ed_get_gimp_matsubara_lattice_21(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: s1, s2, a, b
    if(.not.allocated(Gmatsii))stop "ed_get_gimp_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Gmatsii,1))stop "ed_get_gimp_matsubara error: site index out of range"
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb, b+(s2-1)*Norb, :) = Gmatsii(ilat,s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_matsubara_lattice_21

  subroutine ed_get_gimp_matsubara_lattice_31(F,ilat,ispin,jspin,iorb,jorb)
    integer :: ilat, ispin, jspin, iorb, jorb
    complex(8),dimension(Lmats),intent(inout) :: F
    if(.not.allocated(Gmatsii))stop "ed_get_gimp_matsubara error: no lattice solve"
    if(ilat<1.or.ilat>size(Gmatsii,1))stop "ed_get_gimp_matsubara error: site index out of range"
    F = Gmatsii(ilat,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_matsubara_lattice_31

  subroutine ed_get_gimp_real_lattice_11(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: F
    if(.not.allocated(Grealii))stop "ed_get_gimp_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Grealii,1))stop "ed_get_gimp_real error: site index out of range"
    F = Grealii(ilat,:,:,:,:,:)
  end subroutine ed_get_gimp_real_lattice_11

  subroutine ed_get_gimp_real_lattice_21(F,ilat)
    integer :: ilat
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: s1, s2, a, b
    if(.not.allocated(Grealii))stop "ed_get_gimp_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Grealii,1))stop "ed_get_gimp_real error: site index out of range"
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb, b+(s2-1)*Norb, :) = Grealii(ilat,s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_real_lattice_21

  subroutine ed_get_gimp_real_lattice_31(F,ilat,ispin,jspin,iorb,jorb)
    integer :: ilat, ispin, jspin, iorb, jorb
    complex(8),dimension(Lreal),intent(inout) :: F
    if(.not.allocated(Grealii))stop "ed_get_gimp_real error: no lattice solve"
    if(ilat<1.or.ilat>size(Grealii,1))stop "ed_get_gimp_real error: site index out of range"
    F = Grealii(ilat,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_real_lattice_31

  subroutine anomalous_zero_lattice_11(F,ilat)
    complex(8),intent(inout) :: F(:,:,:,:,:)
    integer :: ilat
    F = (0d0,0d0)
  end subroutine anomalous_zero_lattice_11

  subroutine anomalous_zero_lattice_21(F,ilat)
    complex(8),intent(inout) :: F(:,:,:)
    integer :: ilat
    F = (0d0,0d0)
  end subroutine anomalous_zero_lattice_21

  !> ed_get_phisc (ED_IO/get_phisc.f90): the superconducting order parameter
  !> per orbital, zero in the normal mode
  subroutine ed_get_phisc_1(phisc)
    real(8),dimension(Norb) :: phisc
    phisc = 0d0
  end subroutine ed_get_phisc_1

  subroutine ed_get_phisc_2(phisc,iorb)
    real(8) :: phisc
    integer :: iorb
    if(iorb>Norb)stop "ed_get_phisc error: orbital index > N_orbital"
    phisc = 0d0
  end subroutine ed_get_phisc_2

  subroutine ed_get_phisc_lattice_1(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,Norb) :: yii
    yii = 0d0
  end subroutine ed_get_phisc_lattice_1

  subroutine ed_get_phisc_lattice_2(yii,Nlat,iorb)
    integer :: Nlat, iorb
    real(8),dimension(Nlat) :: yii
    if(iorb>Norb)stop "ed_get_phisc error: orbital index > N_orbital"
    yii = 0d0
  end subroutine ed_get_phisc_lattice_2

  !> so2os_reshape / os2so_reshape (ED_AUX_FUNX.f90:354-404): the
  !> (Nspin*Norb)^2 matrix between the spin-major index io = a + (s-1)*Norb
  !> and the orbital-major index io = s + (a-1)*Nspin
  function so2os_reshape(fg,Nspin_,Norb_) result(g)
    integer    :: Nspin_, Norb_
    complex(8) :: fg(Nspin_*Norb_,Nspin_*Norb_), g(Nspin_*Norb_,Nspin_*Norb_)
    integer    :: a, b, s1, s2
    do s1=1,Nspin_
       do s2=1,Nspin_
          do a=1,Norb_
             do b=1,Norb_
                g(a+(s1-1)*Norb_, b+(s2-1)*Norb_) = fg(s1+(a-1)*Nspin_, s2+(b-1)*Nspin_)
             enddo
          enddo
       enddo
    enddo
  end function so2os_reshape

  function os2so_reshape(fg,Nspin_,Norb_) result(g)
    integer    :: Nspin_, Norb_
    complex(8) :: fg(Nspin_*Norb_,Nspin_*Norb_), g(Nspin_*Norb_,Nspin_*Norb_)
    integer    :: a, b, s1, s2
    do s1=1,Nspin_
       do s2=1,Nspin_
          do a=1,Norb_
             do b=1,Norb_
                g(s1+(a-1)*Nspin_, s2+(b-1)*Nspin_) = fg(a+(s1-1)*Norb_, b+(s2-1)*Norb_)
             enddo
          enddo
       enddo
    enddo
  end function os2so_reshape

  !---------------------------------------------------------------- bath components
  !> get/set/copy_component_bath and the component dimensions
  !> (ED_BATH/user_aux.f90:96-168, 185-334, 431-583, 678-894; type check
  !> ED_BATH.f90:165-176). itype 1 is the bath levels e, itype 2 the
  !> hybridisations v. The hybrid bath has one level set per spin,
  !> e(Nspin,1,Nbath); its v is (Nspin,Norb,Nbath) as for the normal bath.
  subroutine check_type_bath(itype)
    integer :: itype
    if(itype<1.or.itype>2)stop "check_type_bath error: ed_mode=normal, itype!=1,2"
  end subroutine check_type_bath

  pure integer function bath_levels_per_spin()
    bath_levels_per_spin = Norb
    if(trim(bath_type)=='hybrid')bath_levels_per_spin = 1
  end function bath_levels_per_spin

  function get_component_bath_dimension(itype) result(Ndim)
    integer :: itype, Ndim(3)
    call check_type_bath(itype)
    Ndim = [Nspin,Norb,Nbath]
    if(itype==1)Ndim(2) = bath_levels_per_spin()
  end function get_component_bath_dimension

  function get_spin_component_bath_dimension(itype) result(Ndim)
    integer :: itype, Ndim(2)
    call check_type_bath(itype)
    Ndim = [Norb,Nbath]
    if(itype==1)Ndim(1) = bath_levels_per_spin()
  end function get_spin_component_bath_dimension

  function get_orb_component_bath_dimension(itype) result(Ndim)
    integer :: itype, Ndim(2)
    Ndim = [Nspin,Nbath]
  end function get_orb_component_bath_dimension

  function get_spin_orb_component_bath_dimension(itype) result(Ndim)
    integer :: itype, Ndim
    Ndim = Nbath
  end function get_spin_orb_component_bath_dimension

  !> flat bath vector -> e(Nspin,levels,Nbath), v(Nspin,Norb,Nbath) and back,
  !> in the layout of set_bath/get_bath, on arrays of the caller's
  subroutine bath_split(bath,e,v,who)
    real(8),intent(in)  :: bath(:)
    real(8),intent(out) :: e(:,:,:), v(:,:,:)
    character(len=*)    :: who
    integer :: s, a, k, ne, st
    if(size(bath)/=get_bath_dimension())stop trim(who)//" error: wrong bath dimensions"
    ne = bath_levels_per_spin()
    st = Nspin*ne*Nbath
    do s=1,Nspin
       do a=1,ne
          do k=1,Nbath
             e(s,a,k) = bath(k + (a-1)*Nbath + (s-1)*ne*Nbath)
          enddo
       enddo
       do a=1,Norb
          do k=1,Nbath
             v(s,a,k) = bath(st + k + (a-1)*Nbath + (s-1)*Norb*Nbath)
          enddo
       enddo
    enddo
  end subroutine bath_split

  subroutine bath_join(e,v,bath)
    real(8),intent(in)    :: e(:,:,:), v(:,:,:)
    real(8),intent(inout) :: bath(:)
    integer :: s, a, k, ne, st
    ne = bath_levels_per_spin()
    st = Nspin*ne*Nbath
    do s=1,Nspin
       do a=1,ne
          do k=1,Nbath
             bath(k + (a-1)*Nbath + (s-1)*ne*Nbath) = e(s,a,k)
          enddo
       enddo
       do a=1,Norb
          do k=1,Nbath
             bath(st + k + (a-1)*Nbath + (s-1)*Norb*Nbath) = v(s,a,k)
          enddo
       enddo
    enddo
  end subroutine bath_join

  subroutine get_full_component_bath(array,bath_,itype)
    real(8),dimension(:,:,:) :: array
    real(8),dimension(:)     :: bath_
    integer                  :: itype
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"get_component_bath")
    if(any(shape(array)/=get_component_bath_dimension(itype)))&
         stop "get_component_bath error: Array has the wrong shape"
    if(itype==1)then
       array = e
    else
       array = v
    endif
  end subroutine get_full_component_bath

  subroutine get_spin_component_bath(array,bath_,itype,ispin)
    real(8),dimension(:,:) :: array
    real(8),dimension(:)   :: bath_
    integer                :: itype, ispin
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"get_spin_component_bath")
    if(any(shape(array)/=get_spin_component_bath_dimension(itype)))&
         stop "get_spin_component_bath error: Array has the wrong shape"
    if(ispin>Nspin)stop "get_spin_component_bath error: ispin > Nspin"
    if(itype==1)then
       array = e(ispin,:,:)
    else
       array = v(ispin,:,:)
    endif
  end subroutine get_spin_component_bath

  subroutine get_spin_orb_component_bath(array,bath_,itype,ispin,iorb)
    real(8),dimension(:) :: array
    real(8),dimension(:) :: bath_
    integer              :: itype, ispin, iorb
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"get_spin_orb_component_bath")
    call check_type_bath(itype)
    if(size(array)/=Nbath)stop "get_spin_orb_component_bath error: Array has the wrong shape"
    if(ispin>Nspin)stop "get_spin_orb_component_bath error: ispin > Nspin"
    if(iorb>Norb)stop "get_spin_orb_component_bath error: iorb > Norb"
    if(itype==1)then
       array = e(ispin,min(iorb,bath_levels_per_spin()),:)
    else
       array = v(ispin,iorb,:)
    endif
  end subroutine get_spin_orb_component_bath

  subroutine set_full_component_bath(array,bath_,itype)
    real(8),dimension(:,:,:) :: array
    real(8),dimension(:)     :: bath_
    integer                  :: itype
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"set_component_bath")
    if(any(shape(array)/=get_component_bath_dimension(itype)))&
         stop "set_component_bath error: Array has the wrong shape"
    if(itype==1)then
       e = array
    else
       v = array
    endif
    call bath_join(e,v,bath_)
  end subroutine set_full_component_bath

  subroutine set_spin_component_bath(array,bath_,itype,ispin)
    real(8),dimension(:,:) :: array
    real(8),dimension(:)   :: bath_
    integer                :: itype, ispin
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"set_spin_component_bath")
    if(any(shape(array)/=get_spin_component_bath_dimension(itype)))&
         stop "set_spin_component_bath error: Array has the wrong shape"
    if(ispin>Nspin)stop "set_spin_component_bath error: ispin > Nspin"
    if(itype==1)then
       e(ispin,:,:) = array
    else
       v(ispin,:,:) = array
    endif
    call bath_join(e,v,bath_)
  end subroutine set_spin_component_bath

  subroutine set_spin_orb_component_bath(array,bath_,itype,ispin,iorb)
    real(8),dimension(:) :: array
    real(8),dimension(:) :: bath_
    integer              :: itype, ispin, iorb
    real(8) :: e(Nspin,bath_levels_per_spin(),Nbath), v(Nspin,Norb,Nbath)
    call bath_split(bath_,e,v,"set_spin_orb_component_bath")
    call check_type_bath(itype)
    if(size(array)/=Nbath)stop "set_spin_orb_component_bath error: Array has the wrong shape"
    if(ispin>Nspin)stop "set_spin_orb_component_bath error: ispin > Nspin"
    if(iorb>Norb)stop "set_spin_orb_component_bath error: iorb > Norb"
    if(itype==1)then
       e(ispin,min(iorb,bath_levels_per_spin()),:) = array
    else
       v(ispin,iorb,:) = array
    endif
    call bath_join(e,v,bath_)
  end subroutine set_spin_orb_component_bath

  subroutine copy_full_component_bath(bathIN,bathOUT,itype)
    real(8),dimension(:) :: bathIN, bathOUT
    integer              :: itype
    real(8) :: ei(Nspin,bath_levels_per_spin(),Nbath), vi(Nspin,Norb,Nbath)
    real(8) :: eo(Nspin,bath_levels_per_spin(),Nbath), vo(Nspin,Norb,Nbath)
    call bath_split(bathIN,ei,vi,"copy_component_bath")
    call bath_split(bathOUT,eo,vo,"copy_component_bath")
    call check_type_bath(itype)
    if(itype==1)then
       eo = ei
    else
       vo = vi
    endif
    call bath_join(eo,vo,bathOUT)
  end subroutine copy_full_component_bath

  subroutine copy_spin_component_bath(bathIN,ispin,bathOUT,jspin,itype)
    real(8),dimension(:) :: bathIN, bathOUT
    integer              :: ispin, jspin
    integer,optional     :: itype
    real(8) :: ei(Nspin,bath_levels_per_spin(),Nbath), vi(Nspin,Norb,Nbath)
    real(8) :: eo(Nspin,bath_levels_per_spin(),Nbath), vo(Nspin,Norb,Nbath)
    call bath_split(bathIN,ei,vi,"copy_component_bath")
    call bath_split(bathOUT,eo,vo,"copy_component_bath")
    if(present(itype))call check_type_bath(itype)
    if(ispin>Nspin.or.jspin>Nspin)stop "copy_spin_component_bath error: ispin/jspin > Nspin"
    if(.not.present(itype))then
       eo(jspin,:,:) = ei(ispin,:,:)
       vo(jspin,:,:) = vi(ispin,:,:)
    elseif(itype==1)then
       eo(jspin,:,:) = ei(ispin,:,:)
    else
       vo(jspin,:,:) = vi(ispin,:,:)
    endif
    call bath_join(eo,vo,bathOUT)
  end subroutine copy_spin_component_bath

  subroutine copy_spin_orb_component_bath(bathIN,ispin,iorb,bathOUT,jspin,jorb,itype)
    real(8),dimension(:) :: bathIN, bathOUT
    integer              :: ispin, jspin, iorb, jorb
    integer,optional     :: itype
    real(8) :: ei(Nspin,bath_levels_per_spin(),Nbath), vi(Nspin,Norb,Nbath)
    real(8) :: eo(Nspin,bath_levels_per_spin(),Nbath), vo(Nspin,Norb,Nbath)
    integer :: ai, ao
    call bath_split(bathIN,ei,vi,"copy_spin_orb_component_bath")
    call bath_split(bathOUT,eo,vo,"copy_spin_orb_component_bath")
    if(present(itype))call check_type_bath(itype)
    if(ispin>Nspin.or.jspin>Nspin)stop "copy_spin_orb_component_bath error: ispin/jspin > Nspin"
    if(iorb>Norb.or.jorb>Norb)stop "copy_spin_orb_component_bath error: iorb/jorb > Norb"
    ai = min(iorb,bath_levels_per_spin())
    ao = min(jorb,bath_levels_per_spin())
    if(.not.present(itype))then
       eo(jspin,ao,:) = ei(ispin,ai,:)
       vo(jspin,jorb,:) = vi(ispin,iorb,:)
    elseif(itype==1)then
       eo(jspin,ao,:) = ei(ispin,ai,:)
    else
       vo(jspin,jorb,:) = vi(ispin,iorb,:)
    endif
    call bath_join(eo,vo,bathOUT)
  end subroutine copy_spin_orb_component_bath

  !> ed_init_solver_single (ED_MAIN.f90:67-107)
  subroutine ed_init_solver_single(bath,Hloc)
    real(8),dimension(:),intent(inout) :: bath
    complex(8),intent(in)              :: Hloc(Nspin,Nspin,Norb,Norb)
    integer :: Ns
    call check_modes()
    if(size(bath)/=get_bath_dimension())stop "init_ed_solver_single error: wrong bath dimensions"
    if(isetup)then
       call ed_check(dmed_init(0_c_int),"dmed_init")
       if(allocated(impHloc))deallocate(impHloc)
       allocate(impHloc(Nspin,Nspin,Norb,Norb)); impHloc=0d0
       allocate(impGmats(Nspin,Nspin,Norb,Norb,Lmats), impSmats(Nspin,Nspin,Norb,Norb,Lmats))
       allocate(impGreal(Nspin,Nspin,Norb,Norb,Lreal), impSreal(Nspin,Nspin,Norb,Norb,Lreal))
       allocate(ed_dens(Norb), ed_docc(Norb))
       if(allocated(ed_dens_up))deallocate(ed_dens_up)
       if(allocated(ed_dens_dw))deallocate(ed_dens_dw)
       allocate(ed_dens_up(Norb), ed_dens_dw(Norb)); ed_dens_up = 0d0; ed_dens_dw = 0d0
       allocate(bath_e(Nspin,Norb,Nbath), bath_v(Nspin,Norb,Nbath))
       Ns = (Nbath+1)*Norb
       if(trim(bath_type)=='hybrid')Ns = Nbath+Norb
       allocate(neigen_sector((Ns+1)**2)); neigen_sector = 0
       if(lanc_nstates_total/=1)then
          if(mod(lanc_nstates_sector,2)/=0)lanc_nstates_sector=lanc_nstates_sector+1
          if(mod(lanc_nstates_total,2)/=0)lanc_nstates_total=lanc_nstates_total+1
       endif
    endif
    call set_hloc(Hloc)
    bath = 0d0
    call init_bath()
    call get_bath(bath)
    isetup=.false.
  end subroutine ed_init_solver_single

  !> dmed_solve_params from the input variables
  subroutine fill_params(p)
    type(dmed_solve_params_t),intent(out) :: p
    call dmed_solve_params_default(p)
    p%lanc_nstates_sector = lanc_nstates_sector
    p%lanc_nstates_total  = lanc_nstates_total
    p%lanc_nstates_step   = lanc_nstates_step
    p%lanc_niter   = lanc_niter
    p%lanc_ngfiter = lanc_ngfiter
    p%lanc_ncv_factor = lanc_ncv_factor
    p%lanc_ncv_add = lanc_ncv_add
    p%lanc_dim_threshold = lanc_dim_threshold
    p%lanc_tolerance = lanc_tolerance
    p%gs_threshold = gs_threshold
    p%cutoff = cutoff
    p%beta = beta; p%eps = eps; p%wini = wini; p%wfin = wfin
    p%lmats = Lmats; p%lreal = Lreal
    p%prec_gf = merge(DMED_MIXED,DMED_FP64,ed_gpu_fp32_gf)
    p%batched = merge(1,0,ed_gpu_batched)
    p%prec_gs = merge(DMED_MIXED,DMED_FP64,ed_gpu_mixed_gs)
    p%verbose = ed_verbose
    p%chi_spin = merge(1,0,chiflag)
    p%chi_dens = merge(1,0,chiflag)
    p%ltau = Ltau
  end subroutine fill_params

  !> The model this module holds -- the bath last set by set_bath (or
  !> ed_solve), impHloc and the interaction parameters -- handed to the
  !> library, which builds every sector's tables from it.  This is the state
  !> the reference's H*v reads from module variables (dmft_bath, impHloc,
  !> Uloc, Ust, Jh, xmu, ...; ED_HAMILTONIAN_MATVEC.f90:60-100).
  subroutine ed_push_model()
    type(dmed_model_t)         :: m
    real(8),allocatable,target :: bcopy(:), hcopy(:,:,:,:)
    allocate(bcopy(get_bath_dimension()))
    call get_bath(bcopy)
    allocate(hcopy(Nspin,Nspin,Norb,Norb)); hcopy = impHloc
    m%norb = Norb; m%nbath = Nbath; m%nspin = Nspin
    m%bath = c_loc(bcopy); m%hloc = c_loc(hcopy)
    m%uloc = Uloc; m%ust = Ust; m%jh = Jh; m%jx = Jx; m%jp = Jp; m%xmu = xmu
    m%hfmode = merge(1,0,HFmode); m%jhflag = merge(1,0,Jhflag)
    m%bath_type = merge(DMED_BATH_HYBRID,DMED_BATH_NORMAL,trim(bath_type)=='hybrid')
    call ed_check(dmed_set_model(m),"dmed_set_model")
  end subroutine ed_push_model

  !> ed_solve_single (ED_MAIN.f90:251-289)
  subroutine ed_solve_single(bath,Hloc)
    real(8),dimension(:),intent(in),target :: bath
    complex(8),optional,intent(in)         :: Hloc(Nspin,Nspin,Norb,Norb)
    type(dmed_solve_params_t) :: p
    type(dmed_solve_out_t)    :: o
    real(8),allocatable,target :: dens(:), docc(:)
    complex(8),allocatable,target :: gm(:,:,:,:,:), gr(:,:,:,:,:), sm(:,:,:,:,:), sr(:,:,:,:,:)
    integer(c_int32_t),target :: ntot
    integer                   :: lt
    if(isetup)stop "ed_solve: call ed_init_solver first"
    if(present(Hloc))call set_hloc(Hloc)
    if(size(bath)/=get_bath_dimension())stop "ED_SOLVE_SINGLE Error: wrong bath dimensions"
    call set_bath(bath)
    call save_bath(used=.true.)
    call ed_push_model()
    call fill_params(p)
    allocate(dens(Norb), docc(Norb))
    allocate(gm(Nspin,Nspin,Norb,Norb,Lmats), sm(Nspin,Nspin,Norb,Norb,Lmats))
    allocate(gr(Nspin,Nspin,Norb,Norb,Lreal), sr(Nspin,Nspin,Norb,Norb,Lreal))
    ntot = lanc_nstates_total
    o%dens = c_loc(dens); o%docc = c_loc(docc)
    o%gmats = c_loc(gm); o%greal = c_loc(gr); o%smats = c_loc(sm); o%sreal = c_loc(sr)
    o%neigen_sector = c_loc(neigen_sector)
    o%lanc_nstates_total = c_loc(ntot)
    o%obs = c_loc(ed_obs)
    if(chiflag)then
       ! Ltau=max(int(beta),Ltau) (ED_INPUT_VARS.f90:191)
       if(allocated(spinChi_tau))deallocate(spinChi_tau,spinChi_iv,spinChi_w)
       allocate(spinChi_tau(Norb+1,0:max(int(beta),Ltau)))
       allocate(spinChi_iv(Norb+1,0:Lmats), spinChi_w(Norb+1,Lreal))
       o%spin_chi_tau = c_loc(spinChi_tau)
       o%spin_chi_iv = c_loc(spinChi_iv)
       o%spin_chi_w = c_loc(spinChi_w)
       lt = max(int(beta),Ltau)
       if(allocated(densChi_tau))deallocate(densChi_tau,densChi_iv,densChi_w,densChi_mix_tau,&
            densChi_mix_iv,densChi_mix_w,densChi_tot_tau,densChi_tot_iv,densChi_tot_w)
       allocate(densChi_tau(Norb,Norb,0:lt), densChi_mix_tau(Norb,Norb,0:lt))
       allocate(densChi_iv(Norb,Norb,0:Lmats), densChi_mix_iv(Norb,Norb,0:Lmats))
       allocate(densChi_w(Norb,Norb,Lreal), densChi_mix_w(Norb,Norb,Lreal))
       allocate(densChi_tot_tau(0:lt), densChi_tot_iv(0:Lmats), densChi_tot_w(Lreal))
       o%dens_chi_tau = c_loc(densChi_tau); o%dens_chi_iv = c_loc(densChi_iv)
       o%dens_chi_w = c_loc(densChi_w)
       o%dens_chi_mix_tau = c_loc(densChi_mix_tau); o%dens_chi_mix_iv = c_loc(densChi_mix_iv)
       o%dens_chi_mix_w = c_loc(densChi_mix_w)
       o%dens_chi_tot_tau = c_loc(densChi_tot_tau); o%dens_chi_tot_iv = c_loc(densChi_tot_iv)
       o%dens_chi_tot_w = c_loc(densChi_tot_w)
    endif
    call ed_check(dmed_solve(p,o),"dmed_solve")
    lanc_nstates_total = ntot
    impGmats = gm; impGreal = gr; impSmats = sm; impSreal = sr
    ed_dens = dens; ed_docc = docc; ed_egs = o%egs
    ed_dens_up = ed_obs%dens_up(1:Norb); ed_dens_dw = ed_obs%dens_dw(1:Norb)
    ed_Epot = ed_obs%epot; ed_Eint = ed_obs%eint; ed_Ehartree = ed_obs%ehartree
    ed_Eknot = ed_obs%eknot
    ed_Dust = ed_obs%dust; ed_Dund = ed_obs%dund; ed_Dse = ed_obs%dse; ed_Dph = ed_obs%dph
    if(ed_verbose>=1)then
       write(LOGfile,"(A,F20.12,A,I4,A,F10.3,A,F10.3,A)")"Egs =",o%egs,"  states:",o%nstates,&
            "  t_diag=",o%t_diag,"s  t_gf=",o%t_gf,"s"
       write(LOGfile,"(A,10F18.12)")"dens=",ed_dens
       write(LOGfile,"(A,10F18.12)")"docc=",ed_docc
    endif
    if(ed_print_Sigma)call ed_print_impSigma()
    if(ed_print_G)call ed_print_impG()
    if(ed_print_G0)call ed_print_impG0()
    if(chiflag)then
       call print_chi_spin()
       call print_chi_dens()
    endif
  end subroutine ed_solve_single

  !> suffix of site ilat: ineq_site_suffix "_ineq" and site_indx_padding 4
  !> (ED_VARS_GLOBAL.f90:214-215; ED_MAIN.f90:185)
  function site_suffix(ilat) result(sfx)
    integer,intent(in) :: ilat
    character(len=16)  :: sfx
    write(sfx,"(A,I4.4)")"_ineq",ilat
  end function site_suffix

  !> ed_init_solver_lattice (ED_MAIN.f90:170-195): every site initialised as a
  !> single site (its bath from hamiltonian_ineq<site>.restart if present)
  subroutine ed_init_solver_lattice(bath,Hloc)
    real(8),dimension(:,:),intent(inout) :: bath
    complex(8),intent(in)                :: Hloc(size(bath,1),Nspin,Nspin,Norb,Norb)
    integer :: ilat, Nlat, Ns
    Nlat = size(bath,1)
    do ilat=1,Nlat
       ed_file_suffix = site_suffix(ilat)
       call ed_init_solver_single(bath(ilat,:),Hloc(ilat,:,:,:,:))
    enddo
    ed_file_suffix = ""
    Ns = (Nbath+1)*Norb
    if(trim(bath_type)=='hybrid')Ns = Nbath+Norb
    if(allocated(neigen_sectorii))deallocate(neigen_sectorii,neigen_totalii)
    allocate(neigen_sectorii((Ns+1)**2,Nlat), neigen_totalii(Nlat))
    neigen_sectorii = 0
    neigen_totalii = lanc_nstates_total
  end subroutine ed_init_solver_lattice

  !> ed_solve_lattice (ED_MAIN.f90:364-468): one impurity problem per site,
  !> optional per-site Uloc, Ust, Jh; all sites in one dmed_solve_sites call,
  !> which runs them concurrently on the GPU.  Per site the bath is saved to
  !> hamiltonian_ineq<site>.used and Sigma to impSigma_..._ineq<site>.ed.
  subroutine ed_solve_lattice(bath,Hloc,Uloc_ii,Ust_ii,Jh_ii)
    real(8),intent(in)           :: bath(:,:)
    complex(8),intent(in)        :: Hloc(size(bath,1),Nspin,Nspin,Norb,Norb)
    real(8),optional,intent(in)  :: Uloc_ii(size(bath,1),Norb)
    real(8),optional,intent(in)  :: Ust_ii(size(bath,1))
    real(8),optional,intent(in)  :: Jh_ii(size(bath,1))
    integer :: Nlat, ilat, Nb
    type(dmed_model_t),allocatable        :: m(:)
    type(dmed_solve_out_t),allocatable    :: o(:)
    type(dmed_solve_params_t)             :: p
    type(dmed_local_obs_t),allocatable,target :: obs(:)
    real(8),allocatable,target    :: bcopy(:,:), hcopy(:,:,:,:,:), dens(:,:), docc(:,:)
    complex(8),allocatable,target :: gm(:,:,:,:,:,:), gr(:,:,:,:,:,:), sm(:,:,:,:,:,:), sr(:,:,:,:,:,:)
    ! per-site susceptibilities (CHIFLAG), site index last
    real(8),allocatable,target    :: cst(:,:,:), cdt(:,:,:,:), cmt(:,:,:,:), ctt(:,:)
    complex(8),allocatable,target :: csi(:,:,:), csw(:,:,:), cdi(:,:,:,:), cdw(:,:,:,:)
    complex(8),allocatable,target :: cmi(:,:,:,:), cmw(:,:,:,:), cti(:,:), ctw(:,:)
    integer :: lt
    if(isetup)stop "ed_solve_lattice: call ed_init_solver first"
    if(.not.allocated(neigen_sectorii))stop "ed_solve_lattice: call ed_init_solver with the lattice bath first"
    Nlat = size(bath,1)
    Nb = get_bath_dimension()
    if(size(bath,2)/=Nb)stop "init_lattice_bath: wrong bath size dimension 1 or 2 "
    if(size(neigen_totalii)<Nlat)stop "ed_solve_lattice error: size(neigen_totalii,1)<Nsites"
    allocate(m(Nlat), o(Nlat), obs(Nlat))
    allocate(bcopy(Nb,Nlat), hcopy(Nspin,Nspin,Norb,Norb,Nlat), dens(Norb,Nlat), docc(Norb,Nlat))
    allocate(gm(Nspin,Nspin,Norb,Norb,Lmats,Nlat), sm(Nspin,Nspin,Norb,Norb,Lmats,Nlat))
    allocate(gr(Nspin,Nspin,Norb,Norb,Lreal,Nlat), sr(Nspin,Nspin,Norb,Norb,Lreal,Nlat))
    call fill_params(p)
    p%batched = 1
    do ilat=1,Nlat
       if(any(abs(aimag(Hloc(ilat,:,:,:,:)))>0d0))stop "DMFT_ED: complex Hloc is not supported on the GPU path (H must be real)"
       bcopy(:,ilat) = bath(ilat,:)
       hcopy(:,:,:,:,ilat) = dble(Hloc(ilat,:,:,:,:))
       m(ilat)%norb = Norb; m(ilat)%nbath = Nbath; m(ilat)%nspin = Nspin
       m(ilat)%bath = c_loc(bcopy(1,ilat)); m(ilat)%hloc = c_loc(hcopy(1,1,1,1,ilat))
       m(ilat)%uloc = Uloc; m(ilat)%ust = Ust; m(ilat)%jh = Jh
       if(present(Uloc_ii))m(ilat)%uloc(1:Norb) = Uloc_ii(ilat,1:Norb)
       if(present(Ust_ii))m(ilat)%ust = Ust_ii(ilat)
       if(present(Jh_ii))m(ilat)%jh = Jh_ii(ilat)
       m(ilat)%jx = Jx; m(ilat)%jp = Jp; m(ilat)%xmu = xmu
       m(ilat)%hfmode = merge(1,0,HFmode); m(ilat)%jhflag = merge(1,0,Jhflag)
       m(ilat)%bath_type = merge(DMED_BATH_HYBRID,DMED_BATH_NORMAL,trim(bath_type)=='hybrid')
       o(ilat)%dens = c_loc(dens(1,ilat)); o(ilat)%docc = c_loc(docc(1,ilat))
       o(ilat)%gmats = c_loc(gm(1,1,1,1,1,ilat)); o(ilat)%greal = c_loc(gr(1,1,1,1,1,ilat))
       o(ilat)%smats = c_loc(sm(1,1,1,1,1,ilat)); o(ilat)%sreal = c_loc(sr(1,1,1,1,1,ilat))
       o(ilat)%neigen_sector = c_loc(neigen_sectorii(1,ilat))
       o(ilat)%lanc_nstates_total = c_loc(neigen_totalii(ilat))
       o(ilat)%obs = c_loc(obs(ilat))
       o(ilat)%spin_chi_iv = c_null_ptr; o(ilat)%spin_chi_tau = c_null_ptr; o(ilat)%spin_chi_w = c_null_ptr
       o(ilat)%dens_chi_iv = c_null_ptr; o(ilat)%dens_chi_tau = c_null_ptr; o(ilat)%dens_chi_w = c_null_ptr
       o(ilat)%dens_chi_mix_iv = c_null_ptr; o(ilat)%dens_chi_mix_tau = c_null_ptr
       o(ilat)%dens_chi_mix_w = c_null_ptr; o(ilat)%dens_chi_tot_iv = c_null_ptr
       o(ilat)%dens_chi_tot_tau = c_null_ptr; o(ilat)%dens_chi_tot_w = c_null_ptr
    enddo
    if(chiflag)then
       lt = max(int(beta),Ltau)
       allocate(cst(Norb+1,0:lt,Nlat), csi(Norb+1,0:Lmats,Nlat), csw(Norb+1,Lreal,Nlat))
       allocate(cdt(Norb,Norb,0:lt,Nlat), cdi(Norb,Norb,0:Lmats,Nlat), cdw(Norb,Norb,Lreal,Nlat))
       allocate(cmt(Norb,Norb,0:lt,Nlat), cmi(Norb,Norb,0:Lmats,Nlat), cmw(Norb,Norb,Lreal,Nlat))
       allocate(ctt(0:lt,Nlat), cti(0:Lmats,Nlat), ctw(Lreal,Nlat))
       do ilat=1,Nlat
          o(ilat)%spin_chi_tau = c_loc(cst(1,0,ilat)); o(ilat)%spin_chi_iv = c_loc(csi(1,0,ilat))
          o(ilat)%spin_chi_w = c_loc(csw(1,1,ilat))
          o(ilat)%dens_chi_tau = c_loc(cdt(1,1,0,ilat)); o(ilat)%dens_chi_iv = c_loc(cdi(1,1,0,ilat))
          o(ilat)%dens_chi_w = c_loc(cdw(1,1,1,ilat))
          o(ilat)%dens_chi_mix_tau = c_loc(cmt(1,1,0,ilat)); o(ilat)%dens_chi_mix_iv = c_loc(cmi(1,1,0,ilat))
          o(ilat)%dens_chi_mix_w = c_loc(cmw(1,1,1,ilat))
          o(ilat)%dens_chi_tot_tau = c_loc(ctt(0,ilat)); o(ilat)%dens_chi_tot_iv = c_loc(cti(0,ilat))
          o(ilat)%dens_chi_tot_w = c_loc(ctw(1,ilat))
       enddo
    endif
    call ed_check(dmed_solve_sites(int(Nlat,c_int32_t),m,p,o),"dmed_solve_sites")
    if(allocated(Smatsii))deallocate(Smatsii,Srealii,Gmatsii,Grealii,nii,dii,mii,eii,ddii)
    allocate(Smatsii(Nlat,Nspin,Nspin,Norb,Norb,Lmats), Srealii(Nlat,Nspin,Nspin,Norb,Norb,Lreal))
    allocate(Gmatsii(Nlat,Nspin,Nspin,Norb,Norb,Lmats), Grealii(Nlat,Nspin,Nspin,Norb,Norb,Lreal))
    allocate(nii(Nlat,Norb), dii(Nlat,Norb), mii(Nlat,Norb), eii(Nlat,4), ddii(Nlat,4))
    do ilat=1,Nlat
       Smatsii(ilat,:,:,:,:,:) = sm(:,:,:,:,:,ilat); Srealii(ilat,:,:,:,:,:) = sr(:,:,:,:,:,ilat)
       Gmatsii(ilat,:,:,:,:,:) = gm(:,:,:,:,:,ilat); Grealii(ilat,:,:,:,:,:) = gr(:,:,:,:,:,ilat)
       nii(ilat,:) = dens(:,ilat); dii(ilat,:) = docc(:,ilat)
       mii(ilat,:) = obs(ilat)%dens_up(1:Norb) - obs(ilat)%dens_dw(1:Norb)
       eii(ilat,:) = [obs(ilat)%epot,obs(ilat)%eint,obs(ilat)%ehartree,obs(ilat)%eknot]
       ddii(ilat,:) = [obs(ilat)%dust,obs(ilat)%dund,obs(ilat)%dse,obs(ilat)%dph]
       ed_file_suffix = site_suffix(ilat)
       call set_bath(bath(ilat,:))
       call set_hloc(Hloc(ilat,:,:,:,:))
       call save_bath(used=.true.)
       if(ed_print_Sigma)call write_gf_files("impSigma",sm(:,:,:,:,:,ilat),sr(:,:,:,:,:,ilat))
       if(ed_print_G)call write_gf_files("impG",gm(:,:,:,:,:,ilat),gr(:,:,:,:,:,ilat))
       if(ed_print_G0)call ed_print_impG0()
       if(chiflag)then
          ! the site's susceptibilities into the module arrays and its files
          ! (buildChi_impurity inside ed_solve_single, ED_MAIN.f90:282)
          if(allocated(spinChi_tau))deallocate(spinChi_tau,spinChi_iv,spinChi_w)
          if(allocated(densChi_tau))deallocate(densChi_tau,densChi_iv,densChi_w,densChi_mix_tau,&
               densChi_mix_iv,densChi_mix_w,densChi_tot_tau,densChi_tot_iv,densChi_tot_w)
          allocate(spinChi_tau(Norb+1,0:lt), spinChi_iv(Norb+1,0:Lmats), spinChi_w(Norb+1,Lreal))
          allocate(densChi_tau(Norb,Norb,0:lt), densChi_mix_tau(Norb,Norb,0:lt))
          allocate(densChi_iv(Norb,Norb,0:Lmats), densChi_mix_iv(Norb,Norb,0:Lmats))
          allocate(densChi_w(Norb,Norb,Lreal), densChi_mix_w(Norb,Norb,Lreal))
          allocate(densChi_tot_tau(0:lt), densChi_tot_iv(0:Lmats), densChi_tot_w(Lreal))
          spinChi_tau = cst(:,:,ilat); spinChi_iv = csi(:,:,ilat); spinChi_w = csw(:,:,ilat)
          densChi_tau = cdt(:,:,:,ilat); densChi_iv = cdi(:,:,:,ilat); densChi_w = cdw(:,:,:,ilat)
          densChi_mix_tau = cmt(:,:,:,ilat); densChi_mix_iv = cmi(:,:,:,ilat)
          densChi_mix_w = cmw(:,:,:,ilat)
          densChi_tot_tau = ctt(:,ilat); densChi_tot_iv = cti(:,ilat); densChi_tot_w = ctw(:,ilat)
          call print_chi_spin()
          call print_chi_dens()
       endif
       if(ed_verbose>=1)write(LOGfile,"(A,I4,A,F20.12,A,I4)")"site",ilat,"  Egs =",o(ilat)%egs,&
            "  states:",o(ilat)%nstates
    enddo
    ed_file_suffix = ""
  end subroutine ed_solve_lattice

  !> <stem>_tau.ed, <stem>_realw.ed, <stem>_iw.ed of one susceptibility
  !> (splot of ED_IO/print_impChi.f90): columns x, y (real) or x, Im y, Re y
  subroutine write_chi_files(stem, ctau, cw, civ)
    character(len=*),intent(in) :: stem
    real(8),intent(in)          :: ctau(0:)
    complex(8),intent(in)       :: cw(:), civ(0:)
    integer :: i, unit, lt
    real(8) :: pi, wr
    pi = acos(-1d0)
    lt = ubound(ctau,1)
    open(newunit=unit,file=stem//"_tau"//trim(ed_file_suffix)//".ed")
    do i=0,lt
       write(unit,"(2ES24.15)")beta*dble(i)/dble(max(lt,1)),ctau(i)
    enddo
    close(unit)
    open(newunit=unit,file=stem//"_realw"//trim(ed_file_suffix)//".ed")
    do i=1,Lreal
       wr = wini
       if(Lreal>1)wr = wini + (wfin-wini)*dble(i-1)/dble(Lreal-1)
       write(unit,"(3ES24.15)")wr,aimag(cw(i)),dble(cw(i))
    enddo
    close(unit)
    open(newunit=unit,file=stem//"_iw"//trim(ed_file_suffix)//".ed")
    do i=0,Lmats
       write(unit,"(3ES24.15)")pi/beta*2d0*dble(i),aimag(civ(i)),dble(civ(i))
    enddo
    close(unit)
  end subroutine write_chi_files

  !> spinChi_l<a>_{tau,realw,iw}.ed and, for Norb > 1, spinChi_tot<Norb+1>_*.ed
  !> (ED_IO/print_impChi.f90:4-18)
  subroutine print_chi_spin()
    integer :: a
    do a=1,Norb
       call write_chi_files("spinChi_l"//itoa(a),spinChi_tau(a,:),spinChi_w(a,:),spinChi_iv(a,:))
    enddo
    if(Norb>1)call write_chi_files("spinChi_tot"//itoa(Norb+1),spinChi_tau(Norb+1,:),&
         spinChi_w(Norb+1,:),spinChi_iv(Norb+1,:))
  end subroutine print_chi_spin

  !> densChi_l<a><b>_*.ed for a <= b, densChi_mix_l<a><b>_*.ed for all a, b
  !> and densChi_tot_*.ed (ED_IO/print_impChi.f90:29-61)
  subroutine print_chi_dens()
    integer :: a, b
    do a=1,Norb
       do b=a,Norb
          call write_chi_files("densChi_l"//itoa(a)//itoa(b),densChi_tau(a,b,:),&
               densChi_w(a,b,:),densChi_iv(a,b,:))
       enddo
    enddo
    do a=1,Norb
       do b=1,Norb
          call write_chi_files("densChi_mix_l"//itoa(a)//itoa(b),densChi_mix_tau(a,b,:),&
               densChi_mix_w(a,b,:),densChi_mix_iv(a,b,:))
       enddo
    enddo
    call write_chi_files("densChi_tot",densChi_tot_tau,densChi_tot_w,densChi_tot_iv)
  end subroutine print_chi_dens

  !> <stem>_l<a><b>_s<s>_iw<suffix>.ed and _realw: columns x, Im F, Re F
  !> (splot), for a = b, or a <= b with the hybrid bath
  !> (ED_IO/print_impSigma.f90:1-50, print_impG.f90:1-51)
  subroutine write_gf_files(stem,Fm,Fr)
    character(len=*),intent(in) :: stem
    complex(8),intent(in)       :: Fm(Nspin,Nspin,Norb,Norb,Lmats), Fr(Nspin,Nspin,Norb,Norb,Lreal)
    integer :: unit, s, a, b, i
    real(8) :: x
    character(len=:),allocatable :: sfx
    do s=1,Nspin
       do a=1,Norb
          do b=a,Norb
             if(b/=a .and. trim(bath_type)/='hybrid')cycle
             sfx = "_l"//itoa(a)//itoa(b)//"_s"//itoa(s)
             open(newunit=unit,file=stem//sfx//"_iw"//trim(ed_file_suffix)//".ed")
             do i=1,Lmats
                x = acos(-1d0)/beta*dble(2*i-1)
                write(unit,"(3ES24.15)")x,aimag(Fm(s,s,a,b,i)),dble(Fm(s,s,a,b,i))
             enddo
             close(unit)
             open(newunit=unit,file=stem//sfx//"_realw"//trim(ed_file_suffix)//".ed")
             do i=1,Lreal
                x = wini
                if(Lreal>1)x = wini + (wfin-wini)*dble(i-1)/dble(Lreal-1)
                write(unit,"(3ES24.15)")x,aimag(Fr(s,s,a,b,i)),dble(Fr(s,s,a,b,i))
             enddo
             close(unit)
          enddo
       enddo
    enddo
  end subroutine write_gf_files

  !> g0and_bath_mats / _real of the current bath and Hloc: [(z + xmu) - Hloc -
  !> Delta]^-1, orbital-diagonal for the normal bath, the full matrix for the
  !> hybrid bath (ED_BATH_FUNCTIONS.f90:1784-1848; Delta :221-310)
  subroutine compute_g0(G0m,G0r)
    complex(8),intent(out) :: G0m(Nspin,Nspin,Norb,Norb,Lmats), G0r(Nspin,Nspin,Norb,Norb,Lreal)
    integer :: s, i
    complex(8) :: z
    G0m = (0d0,0d0); G0r = (0d0,0d0)
    do s=1,Nspin
       do i=1,Lmats
          z = cmplx(0d0,acos(-1d0)/beta*dble(2*i-1),8)
          call g0_at(s,z,G0m(s,s,:,:,i))
       enddo
       do i=1,Lreal
          z = cmplx(wini,eps,8)
          if(Lreal>1)z = cmplx(wini + (wfin-wini)*dble(i-1)/dble(Lreal-1),eps,8)
          call g0_at(s,z,G0r(s,s,:,:,i))
       enddo
    enddo
  end subroutine compute_g0

  subroutine g0_at(s,z,g)
    integer,intent(in)     :: s
    complex(8),intent(in)  :: z
    complex(8),intent(out) :: g(Norb,Norb)
    complex(8) :: m(Norb,Norb)
    integer :: a, b, k
    m = (0d0,0d0)
    do a=1,Norb
       m(a,a) = z + xmu
       do b=1,Norb
          m(a,b) = m(a,b) - impHloc(s,s,a,b)
          if(trim(bath_type)=='hybrid')then
             do k=1,Nbath
                m(a,b) = m(a,b) - bath_v(s,a,k)*bath_v(s,b,k)/(z - bath_e(s,1,k))
             enddo
          elseif(a==b)then
             do k=1,Nbath
                m(a,a) = m(a,a) - bath_v(s,a,k)**2/(z - bath_e(s,a,k))
             enddo
          endif
       enddo
    enddo
    g = (0d0,0d0)
    if(trim(bath_type)=='hybrid')then
       call invert_cmat(m,g)
    else
       do a=1,Norb
          g(a,a) = 1d0/m(a,a)
       enddo
    endif
  end subroutine g0_at

  !> inverse of a small complex matrix (Gauss-Jordan with partial pivoting)
  subroutine invert_cmat(m,inv)
    complex(8),intent(in)  :: m(:,:)
    complex(8),intent(out) :: inv(size(m,1),size(m,1))
    complex(8) :: w(size(m,1),2*size(m,1)), f, row(2*size(m,1))
    integer :: n, c, r, p
    n = size(m,1)
    w = (0d0,0d0)
    w(:,1:n) = m
    do r=1,n
       w(r,n+r) = (1d0,0d0)
    enddo
    do c=1,n
       p = c
       do r=c+1,n
          if(abs(w(r,c))>abs(w(p,c)))p = r
       enddo
       if(abs(w(p,c))==0d0)stop "DMFT_ED: singular matrix in G0"
       if(p/=c)then
          row = w(c,:); w(c,:) = w(p,:); w(p,:) = row
       endif
       w(c,:) = w(c,:)/w(c,c)
       do r=1,n
          if(r==c)cycle
          f = w(r,c)
          w(r,:) = w(r,:) - f*w(c,:)
       enddo
    enddo
    inv = w(:,n+1:2*n)
  end subroutine invert_cmat

  !> ed_print_impSigma, ed_print_impG, ed_print_impG0 (ED_IO.f90:266-297;
  !> ED_GREENS_FUNCTIONS.f90:161-165).  As in the reference, impG0 is written
  !> to the impG_* file names (print_impG0.f90:43-44), so with ED_PRINT_G0=T
  !> those files hold G0 after a solve.
  subroutine ed_print_impSigma()
    call write_gf_files("impSigma",impSmats,impSreal)
  end subroutine ed_print_impSigma

  subroutine ed_print_impG()
    call write_gf_files("impG",impGmats,impGreal)
  end subroutine ed_print_impG

  subroutine ed_print_impG0()
    complex(8),allocatable :: G0m(:,:,:,:,:), G0r(:,:,:,:,:)
    allocate(G0m(Nspin,Nspin,Norb,Norb,Lmats), G0r(Nspin,Nspin,Norb,Norb,Lreal))
    call compute_g0(G0m,G0r)
    call write_gf_files("impG",G0m,G0r)
  end subroutine ed_print_impG0

  !> ed_print_impChi (ED_IO.f90:299-305): the spin and charge files
  subroutine ed_print_impChi()
    if(allocated(spinChi_tau))call print_chi_spin()
    if(allocated(densChi_tau))call print_chi_dens()
  end subroutine ed_print_impChi

  !> ed_read_impSigma (ED_IO/read_impSigma.f90:1-50): impSmats / impSreal from
  !> the impSigma_*_iw / _realw files written by ed_print_impSigma
  subroutine ed_read_impSigma()
    integer :: unit, s, a, b, i, ios
    real(8) :: x, im, re
    character(len=:),allocatable :: sfx
    if(.not.allocated(impSmats))stop "ed_read_impSigma: call ed_init_solver first"
    do s=1,Nspin
       do a=1,Norb
          do b=a,Norb
             if(b/=a .and. trim(bath_type)/='hybrid')cycle
             sfx = "_l"//itoa(a)//itoa(b)//"_s"//itoa(s)
             open(newunit=unit,file="impSigma"//sfx//"_iw"//trim(ed_file_suffix)//".ed",status='old',action='read')
             do i=1,Lmats
                read(unit,*,iostat=ios)x,im,re
                if(ios/=0)exit
                impSmats(s,s,a,b,i) = cmplx(re,im,8)
             enddo
             close(unit)
             open(newunit=unit,file="impSigma"//sfx//"_realw"//trim(ed_file_suffix)//".ed",status='old',action='read')
             do i=1,Lreal
                read(unit,*,iostat=ios)x,im,re
                if(ios/=0)exit
                impSreal(s,s,a,b,i) = cmplx(re,im,8)
             enddo
             close(unit)
          enddo
       enddo
    enddo
  end subroutine ed_read_impSigma

  !------------------------------------------------------------ getters
  subroutine ed_get_sigma_matsubara_1(Smats)
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: Smats
    Smats = impSmats
  end subroutine ed_get_sigma_matsubara_1

  subroutine ed_get_sigma_real_1(Sreal)
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: Sreal
    Sreal = impSreal
  end subroutine ed_get_sigma_real_1

  subroutine ed_get_gimp_matsubara_1(Gmats)
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: Gmats
    Gmats = impGmats
  end subroutine ed_get_gimp_matsubara_1

  subroutine ed_get_gimp_real_1(Greal)
    complex(8),dimension(Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: Greal
    Greal = impGreal
  end subroutine ed_get_gimp_real_1

  subroutine ed_get_dens_1(dens)
    real(8),dimension(Norb) :: dens
    dens = ed_dens
  end subroutine ed_get_dens_1

  subroutine ed_get_docc_1(docc)
    real(8),dimension(Norb) :: docc
    docc = ed_docc
  end subroutine ed_get_docc_1

  !> ED_IO/get_dens.f90, get_mag.f90 (ed_get_mag_1: dens_up - dens_dw)
  subroutine ed_get_dens_up(dens)
    real(8),dimension(Norb) :: dens
    dens = ed_dens_up
  end subroutine ed_get_dens_up

  subroutine ed_get_dens_dw(dens)
    real(8),dimension(Norb) :: dens
    dens = ed_dens_dw
  end subroutine ed_get_dens_dw

  subroutine ed_get_mag_1(mag)
    real(8),dimension(Norb) :: mag
    mag = ed_dens_up - ed_dens_dw
  end subroutine ed_get_mag_1

  !> ED_IO/get_eimp.f90: [Epot, Eint, Ehartree, Eknot]
  subroutine ed_get_eimp_1(eimp)
    real(8),dimension(4) :: eimp
    eimp = [ed_Epot,ed_Eint,ed_Ehartree,ed_Eknot]
  end subroutine ed_get_eimp_1

  subroutine ed_get_epot(e)
    real(8) :: e
    e = ed_Epot
  end subroutine ed_get_epot

  subroutine ed_get_eint(e)
    real(8) :: e
    e = ed_Eint
  end subroutine ed_get_eint

  subroutine ed_get_ehartree(e)
    real(8) :: e
    e = ed_Ehartree
  end subroutine ed_get_ehartree

  subroutine ed_get_eknot(e)
    real(8) :: e
    e = ed_Eknot
  end subroutine ed_get_eknot

  !> ED_IO/get_doubles.f90: [Dust, Dund, Dse, Dph]
  subroutine ed_get_doubles_1(d)
    real(8),dimension(4) :: d
    d = [ed_Dust,ed_Dund,ed_Dse,ed_Dph]
  end subroutine ed_get_doubles_1

  subroutine ed_get_dust(d)
    real(8) :: d
    d = ed_Dust
  end subroutine ed_get_dust

  subroutine ed_get_dund(d)
    real(8) :: d
    d = ed_Dund
  end subroutine ed_get_dund

  subroutine ed_get_dse(d)
    real(8) :: d
    d = ed_Dse
  end subroutine ed_get_dse

  subroutine ed_get_dph(d)
    real(8) :: d
    d = ed_Dph
  end subroutine ed_get_dph

  !------------------------------------------------------------ chi^2 fit
  !> chi2_fitgf_generic_normal -> chi2_fitgf_normal_normal (ED_CHI2FIT.f90:86-138,
  !> fitgf_normal_normal.f90:18-156): per orbital, fit (e_k, V_k) of spin ispin
  !> to fg on the first Lfit Matsubara points, scheme cg_scheme (weiss/delta),
  !> weights cg_weight; then save the bath (hamiltonian.restart).
  subroutine ed_chi2_fitgf_single(fg,bath,ispin)
    complex(8),dimension(:,:,:,:,:) :: fg
    real(8),dimension(:)            :: bath
    integer,optional                :: ispin
    integer                         :: ispin_, a, L
    real(8)                         :: chi2
    integer(c_int32_t)              :: iter, scheme
    real(8),allocatable             :: e(:), v(:)
    complex(8),allocatable          :: g(:)
    ispin_ = 1; if(present(ispin))ispin_ = ispin
    if(size(fg,1)/=Nspin .or. size(fg,3)/=Norb)stop "chi2_fitgf_generic_normal: wrong fg shape"
    call set_bath(bath)
    L = min(Lfit,size(fg,5))
    if(trim(bath_type)=='hybrid')then
       call fit_hybrid_spin(fg,ispin_,L)
       if(Nspin==2 .and. .not.present(ispin))then
          bath_e(2,:,:) = bath_e(1,:,:); bath_v(2,:,:) = bath_v(1,:,:)
       endif
       call save_bath(used=.false.)
       call get_bath(bath)
       return
    endif
    allocate(e(Nbath), v(Nbath), g(L))
    scheme = 0; if(trim(cg_scheme)=='delta')scheme = 1
    do a=1,Norb
       g = fg(ispin_,ispin_,a,a,1:L)
       e = bath_e(ispin_,a,:); v = bath_v(ispin_,a,:)
       call ed_check(dmed_chi2_fit(scheme,int(cg_weight,c_int32_t),g,int(L,c_int32_t),beta,xmu,&
            impHloc(ispin_,ispin_,a,a),int(Nbath,c_int32_t),e,v,int(cg_niter,c_int32_t),cg_ftol,&
            chi2,iter),"dmed_chi2_fit")
       bath_e(ispin_,a,:) = e; bath_v(ispin_,a,:) = v
       write(LOGfile,"(A,ES18.9,A,I5,A)")"chi^2|iter= ",chi2," | ",iter,"  <--  Orb"//itoa(a)//" Spin"//itoa(ispin_)
    enddo
    if(Nspin==2 .and. .not.present(ispin))then
       bath_e(2,:,:) = bath_e(1,:,:); bath_v(2,:,:) = bath_v(1,:,:)
    endif
    call save_bath(used=.false.)
    call get_bath(bath)
  end subroutine ed_chi2_fitgf_single

  !> chi2_fitgf_hybrid_normal (fitgf_hybrid_normal.f90:18-156): all orbital
  !> pairs a <= b of spin ispin at once, through dmed_chi2_fit_hybrid
  subroutine fit_hybrid_spin(fg,ispin,L)
    complex(8),dimension(:,:,:,:,:),intent(in) :: fg
    integer,intent(in)                         :: ispin, L
    complex(8),allocatable :: g(:,:,:)
    real(8),allocatable    :: e(:), v(:), h(:,:)
    real(8)                :: chi2
    integer(c_int32_t)     :: iter, scheme
    integer                :: a
    allocate(g(Norb,Norb,L), e(Nbath), v(Norb*Nbath), h(Norb,Norb))
    g = fg(ispin,ispin,:,:,1:L)
    h = impHloc(ispin,ispin,:,:)
    e = bath_e(ispin,1,:)
    do a=1,Norb
       v((a-1)*Nbath+1:a*Nbath) = bath_v(ispin,a,:)
    enddo
    scheme = 0; if(trim(cg_scheme)=='delta')scheme = 1
    call ed_check(dmed_chi2_fit_hybrid(scheme,int(cg_weight,c_int32_t),g,int(L,c_int32_t),beta,xmu,&
         h,int(Norb,c_int32_t),int(Nbath,c_int32_t),e,v,int(cg_niter,c_int32_t),cg_ftol,chi2,iter),&
         "dmed_chi2_fit_hybrid")
    do a=1,Norb
       bath_e(ispin,a,:) = e
       bath_v(ispin,a,:) = v((a-1)*Nbath+1:a*Nbath)
    enddo
    write(LOGfile,"(A,ES18.9,A,I5,A)")"chi^2|iter= ",chi2," | ",iter,"  <--  all Orbs, Spin"//itoa(ispin)
  end subroutine fit_hybrid_spin

  !> lattice getters (ED_IO/get_sigma_matsubara.f90:34-38 and the *_lattice
  !> forms of get_dens, get_docc, get_mag, get_eimp, get_doubles)
  subroutine ed_get_sigma_matsubara_lattice_1(Smats,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: Smats
    Smats(1:Nsites,:,:,:,:,:) = Smatsii(1:Nsites,:,:,:,:,:)
  end subroutine ed_get_sigma_matsubara_lattice_1

  subroutine ed_get_sigma_real_lattice_1(Sreal,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: Sreal
    Sreal(1:Nsites,:,:,:,:,:) = Srealii(1:Nsites,:,:,:,:,:)
  end subroutine ed_get_sigma_real_lattice_1

  subroutine ed_get_gimp_matsubara_lattice_1(Gmats,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin,Nspin,Norb,Norb,Lmats),intent(inout) :: Gmats
    Gmats(1:Nsites,:,:,:,:,:) = Gmatsii(1:Nsites,:,:,:,:,:)
  end subroutine ed_get_gimp_matsubara_lattice_1

  subroutine ed_get_gimp_real_lattice_1(Greal,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin,Nspin,Norb,Norb,Lreal),intent(inout) :: Greal
    Greal(1:Nsites,:,:,:,:,:) = Grealii(1:Nsites,:,:,:,:,:)
  end subroutine ed_get_gimp_real_lattice_1

  subroutine ed_get_dens_lattice_1(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,Norb) :: yii
    yii = 0d0
    if(allocated(nii))then
       if(Nlat>size(nii,1))stop "ed_get_dens error: required N_sites > evaluated N_sites"
       yii = nii(1:Nlat,:)
    endif
  end subroutine ed_get_dens_lattice_1

  subroutine ed_get_docc_lattice_1(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,Norb) :: yii
    yii = 0d0
    if(allocated(dii))then
       if(Nlat>size(dii,1))stop "ed_get_docc error: required N_sites > evaluated N_sites"
       yii = dii(1:Nlat,:)
    endif
  end subroutine ed_get_docc_lattice_1

  subroutine ed_get_mag_lattice_1(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,Norb) :: yii
    yii = 0d0
    if(allocated(mii))then
       if(Nlat>size(mii,1))stop "ed_get_mag error: required N_sites > evaluated N_sites"
       yii = mii(1:Nlat,:)
    endif
  end subroutine ed_get_mag_lattice_1

  subroutine ed_get_eimp_lattice(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,4) :: yii
    yii = 0d0
    if(allocated(eii))then
       if(Nlat>size(eii,1))stop "ed_get_eimp error: required N_sites > evaluated N_sites"
       yii = eii(1:Nlat,:)
    endif
  end subroutine ed_get_eimp_lattice

  subroutine ed_get_doubles_lattice(yii,Nlat)
    integer :: Nlat
    real(8),dimension(Nlat,4) :: yii
    yii = 0d0
    if(allocated(ddii))then
       if(Nlat>size(ddii,1))stop "ed_get_doubles error: required N_sites > evaluated N_sites"
       yii = ddii(1:Nlat,:)
    endif
  end subroutine ed_get_doubles_lattice

  !> ed_fit_bath_sites_normal (ED_CHI2FIT.f90:413-452): the single-site fit
  !> on every site with its own Hloc, bath file hamiltonian_ineq<site>.restart
  subroutine ed_fit_bath_sites_normal(bath,Delta,Hloc,ispin)
    real(8),intent(inout)    :: bath(:,:)
    complex(8),intent(inout) :: Delta(size(bath,1),Nspin,Nspin,Norb,Norb,Lmats)
    complex(8)               :: Hloc(size(bath,1),Nspin,Nspin,Norb,Norb)
    integer,optional         :: ispin
    integer                  :: ilat
    real(8),allocatable      :: b(:)
    complex(8),allocatable   :: fg(:,:,:,:,:)
    if(present(ispin))then
       if(ispin>Nspin)stop "ed_fit_bath_sites error: required spin index > Nspin"
    endif
    allocate(b(size(bath,2)), fg(Nspin,Nspin,Norb,Norb,Lmats))
    do ilat=1,size(bath,1)
       call set_hloc(Hloc(ilat,:,:,:,:))
       ed_file_suffix = site_suffix(ilat)
       b = bath(ilat,:)
       fg = Delta(ilat,:,:,:,:,:)
       if(present(ispin))then
          call ed_chi2_fitgf_single(fg,b,ispin)
       else
          call ed_chi2_fitgf_single(fg,b)
       endif
       bath(ilat,:) = b
    enddo
    ed_file_suffix = ""
  end subroutine ed_fit_bath_sites_normal

  !> the (Nspin*Norb)^2 and single-element forms of the getters
  !> (ED_IO/get_sigma_matsubara.f90:7-31 and alike; io = iorb + (ispin-1)*Norb)

  subroutine ed_get_sigma_matsubara_2(F)
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: s1, s2, a, b
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb,b+(s2-1)*Norb,:) = impSmats(s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_matsubara_2

  subroutine ed_get_sigma_matsubara_3(F,ispin,jspin,iorb,jorb)
    complex(8),dimension(Lmats),intent(inout) :: F
    integer :: ispin, jspin, iorb, jorb
    F = impSmats(ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_matsubara_3

  subroutine ed_get_sigma_matsubara_lattice_2(F,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: l, s1, s2, a, b
    do l=1,Nsites
       do s1=1,Nspin
          do s2=1,Nspin
             do a=1,Norb
                do b=1,Norb
                   F(l,a+(s1-1)*Norb,b+(s2-1)*Norb,:) = Smatsii(l,s1,s2,a,b,:)
                enddo
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_matsubara_lattice_2

  subroutine ed_get_sigma_matsubara_lattice_3(F,Nsites,ispin,jspin,iorb,jorb)
    integer :: Nsites, ispin, jspin, iorb, jorb
    complex(8),dimension(Nsites,Lmats),intent(inout) :: F
    F(1:Nsites,:) = Smatsii(1:Nsites,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_matsubara_lattice_3

  subroutine ed_get_sigma_real_2(F)
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: s1, s2, a, b
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb,b+(s2-1)*Norb,:) = impSreal(s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_real_2

  subroutine ed_get_sigma_real_3(F,ispin,jspin,iorb,jorb)
    complex(8),dimension(Lreal),intent(inout) :: F
    integer :: ispin, jspin, iorb, jorb
    F = impSreal(ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_real_3

  subroutine ed_get_sigma_real_lattice_2(F,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: l, s1, s2, a, b
    do l=1,Nsites
       do s1=1,Nspin
          do s2=1,Nspin
             do a=1,Norb
                do b=1,Norb
                   F(l,a+(s1-1)*Norb,b+(s2-1)*Norb,:) = Srealii(l,s1,s2,a,b,:)
                enddo
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_sigma_real_lattice_2

  subroutine ed_get_sigma_real_lattice_3(F,Nsites,ispin,jspin,iorb,jorb)
    integer :: Nsites, ispin, jspin, iorb, jorb
    complex(8),dimension(Nsites,Lreal),intent(inout) :: F
    F(1:Nsites,:) = Srealii(1:Nsites,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_sigma_real_lattice_3

  subroutine ed_get_gimp_matsubara_2(F)
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: s1, s2, a, b
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb,b+(s2-1)*Norb,:) = impGmats(s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_matsubara_2

  subroutine ed_get_gimp_matsubara_3(F,ispin,jspin,iorb,jorb)
    complex(8),dimension(Lmats),intent(inout) :: F
    integer :: ispin, jspin, iorb, jorb
    F = impGmats(ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_matsubara_3

  subroutine ed_get_gimp_matsubara_lattice_2(F,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin*Norb,Nspin*Norb,Lmats),intent(inout) :: F
    integer :: l, s1, s2, a, b
    do l=1,Nsites
       do s1=1,Nspin
          do s2=1,Nspin
             do a=1,Norb
                do b=1,Norb
                   F(l,a+(s1-1)*Norb,b+(s2-1)*Norb,:) = Gmatsii(l,s1,s2,a,b,:)
                enddo
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_matsubara_lattice_2

  subroutine ed_get_gimp_matsubara_lattice_3(F,Nsites,ispin,jspin,iorb,jorb)
    integer :: Nsites, ispin, jspin, iorb, jorb
    complex(8),dimension(Nsites,Lmats),intent(inout) :: F
    F(1:Nsites,:) = Gmatsii(1:Nsites,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_matsubara_lattice_3

  subroutine ed_get_gimp_real_2(F)
    complex(8),dimension(Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: s1, s2, a, b
    do s1=1,Nspin
       do s2=1,Nspin
          do a=1,Norb
             do b=1,Norb
                F(a+(s1-1)*Norb,b+(s2-1)*Norb,:) = impGreal(s1,s2,a,b,:)
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_real_2

  subroutine ed_get_gimp_real_3(F,ispin,jspin,iorb,jorb)
    complex(8),dimension(Lreal),intent(inout) :: F
    integer :: ispin, jspin, iorb, jorb
    F = impGreal(ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_real_3

  subroutine ed_get_gimp_real_lattice_2(F,Nsites)
    integer :: Nsites
    complex(8),dimension(Nsites,Nspin*Norb,Nspin*Norb,Lreal),intent(inout) :: F
    integer :: l, s1, s2, a, b
    do l=1,Nsites
       do s1=1,Nspin
          do s2=1,Nspin
             do a=1,Norb
                do b=1,Norb
                   F(l,a+(s1-1)*Norb,b+(s2-1)*Norb,:) = Grealii(l,s1,s2,a,b,:)
                enddo
             enddo
          enddo
       enddo
    enddo
  end subroutine ed_get_gimp_real_lattice_2

  subroutine ed_get_gimp_real_lattice_3(F,Nsites,ispin,jspin,iorb,jorb)
    integer :: Nsites, ispin, jspin, iorb, jorb
    complex(8),dimension(Nsites,Lreal),intent(inout) :: F
    F(1:Nsites,:) = Grealii(1:Nsites,ispin,jspin,iorb,jorb,:)
  end subroutine ed_get_gimp_real_lattice_3

  subroutine ed_get_dens_2(y,iorb)
    real(8) :: y
    integer :: iorb
    if(iorb>Norb)stop "ed_get_dens error: orbital index > N_orbital"
    y = ed_dens(iorb)
  end subroutine ed_get_dens_2

  subroutine ed_get_dens_lattice_2(yii,Nlat,iorb)
    integer :: Nlat, iorb
    real(8),dimension(Nlat) :: yii
    if(iorb>Norb)stop "ed_get_dens error: orbital index > N_orbital"
    yii = 0d0
    if(allocated(nii))then
       if(Nlat>size(nii,1))stop "ed_get_dens error: required N_sites > evaluated N_sites"
       yii = nii(1:Nlat,iorb)
    endif
  end subroutine ed_get_dens_lattice_2

  subroutine ed_get_docc_2(y,iorb)
    real(8) :: y
    integer :: iorb
    if(iorb>Norb)stop "ed_get_docc error: orbital index > N_orbital"
    y = ed_docc(iorb)
  end subroutine ed_get_docc_2

  subroutine ed_get_docc_lattice_2(yii,Nlat,iorb)
    integer :: Nlat, iorb
    real(8),dimension(Nlat) :: yii
    if(iorb>Norb)stop "ed_get_docc error: orbital index > N_orbital"
    yii = 0d0
    if(allocated(dii))then
       if(Nlat>size(dii,1))stop "ed_get_docc error: required N_sites > evaluated N_sites"
       yii = dii(1:Nlat,iorb)
    endif
  end subroutine ed_get_docc_lattice_2

  subroutine ed_get_mag_2(y,iorb)
    real(8) :: y
    integer :: iorb
    if(iorb>Norb)stop "ed_get_mag error: orbital index > N_orbital"
    y = (ed_dens_up(iorb)-ed_dens_dw(iorb))
  end subroutine ed_get_mag_2

  subroutine ed_get_mag_lattice_2(yii,Nlat,iorb)
    integer :: Nlat, iorb
    real(8),dimension(Nlat) :: yii
    if(iorb>Norb)stop "ed_get_mag error: orbital index > N_orbital"
    yii = 0d0
    if(allocated(mii))then
       if(Nlat>size(mii,1))stop "ed_get_mag error: required N_sites > evaluated N_sites"
       yii = mii(1:Nlat,iorb)
    endif
  end subroutine ed_get_mag_lattice_2

end module DMFT_ED
