! trekis_gpu_bridge.f90 -- do_Monte_Carlo on the MI355X, for the reference's
! own driver.
!
! The reference's MAIN (Universal_MC_for_SHI_MAIN.f90) reads its inputs into
! the derived types of Objects.f90, builds its tables and calls
! do_Monte_Carlo (Monte_Carlo.f90:39-92, called at MAIN:271-276).  This module
! provides a do_Monte_Carlo with the same argument list: compiled with the
! reference's Objects.f90 and this library's trekis_gpu_c.f90, a driver swaps
! `use Monte_Carlo, only : do_Monte_Carlo` (MAIN:75) for
! `use trekis_gpu_bridge, only : do_Monte_Carlo` and nothing else changes.
!
! It flattens the target, ion and options into the library's C structs
! (flatten_material, flatten_ion, flatten_options), creates a context, builds
! the tables on the device (the library computes the same tables the driver
! built, from the same descriptors), runs NMC iterations and adds the
! unnormalised sums into the Out_* arrays, which MAIN then divides by NMC
! (MAIN:281-314).  Errors land in Error_message as Save_error_details would
! put them (report).  Tables passed in (SHI_MFP, Total_el_MFPs, aidCS, ...)
! are not read; Out_Eat_dens, Out_theta1 and the field arrays are left as
! they are, as the reference leaves them with fields off.  MPI_param is not
! used: one process drives one GPU.
module trekis_gpu_bridge
   use iso_c_binding
   use Objects
   use trekis_gpu_c
   implicit none
   private
   public :: do_Monte_Carlo, flatten_material, flatten_ion, flatten_options, report
   public :: trk_bridge_seed, trk_bridge_precision, trk_bridge_first_iter, trk_bridge_e_grid_Emax
   public :: trk_bridge_shi_grid_Emax, bridge_dos

   ! Run controls a driver may set before calling do_Monte_Carlo: the seed of
   ! the per-particle Philox streams, the precision path (0 FP64, replays the
   ! oracle bit for bit; 1 mixed) and the number of the first iteration.
   integer(c_int64_t), save :: trk_bridge_seed = int(z'5EED0001', c_int64_t)
   integer(c_int32_t), save :: trk_bridge_precision = 0
   integer(c_int64_t), save :: trk_bridge_first_iter = 0
   ! Tops of the electron and ion energy grids [eV]; 0 takes the reference's
   ! (Analytical_IMFPs.f90:143, 2272).
   real(c_double), save :: trk_bridge_e_grid_Emax = 0.0d0, trk_bridge_shi_grid_Emax = 0.0d0

   ! Contiguous copies of the DOS arrays the C descriptor points into.
   type :: bridge_dos
      real(c_double), dimension(:), allocatable :: E, DOS, int_DOS, k, Eff_m, k_inv, Eff_m_inv
   end type bridge_dos

contains

   ! Target_atoms / Matter / Mat_DOS / CDF_Phonon (Objects.f90:92-130, 208-265)
   ! into trk_material.  The valence band must be the last shell of the first
   ! atom (Lowest_Ip_At = 1, Lowest_Ip_Shl = its last shell), as the library
   ! keeps it there; every shell must use the CDF cross sections (KOCS 1).
   subroutine flatten_material(Target_atoms, Matter, Mat_DOS, CDF_Phonon, Lowest_Ip_At, Lowest_Ip_Shl, m, dos, &
                               ok, why)
      type(Atom), dimension(:), intent(in) :: Target_atoms
      type(Solid), intent(in) :: Matter
      type(Density_of_states), intent(in) :: Mat_DOS
      type(CDF), intent(in) :: CDF_Phonon
      integer, intent(in) :: Lowest_Ip_At, Lowest_Ip_Shl
      type(trk_material), intent(out) :: m
      type(bridge_dos), intent(inout), target :: dos
      logical, intent(out) :: ok
      character(*), intent(out) :: why
      integer :: a, k, s, o, i, nd
      ok = .false.
      why = ''
      m%n_atoms = size(Target_atoms)
      if (m%n_atoms < 1 .or. m%n_atoms > TRK_MAX_ATOMS) then
         why = 'number of atoms out of range'
         return
      endif
      if (Lowest_Ip_At /= 1 .or. Lowest_Ip_Shl /= Target_atoms(1)%N_shl) then
         why = 'the valence band must be the last shell of the first atom'
         return
      endif
      s = 0
      o = 0
      do a = 1, m%n_atoms
         m%Zat(a) = Target_atoms(a)%Zat
         m%Mass(a) = Target_atoms(a)%Mass
         m%Pers(a) = Target_atoms(a)%Pers
         m%n_shells(a) = Target_atoms(a)%N_shl
         do k = 1, Target_atoms(a)%N_shl
            s = s + 1
            if (s > TRK_MAX_SHELLS) then
               why = 'too many shells'
               return
            endif
            m%kocs(s) = 1
            m%Ek(s) = 0.0d0
            ! the reference computes BEB for every KOCS other than 1 (CS:950)
            if (allocated(Target_atoms(a)%KOCS)) then
               if (Target_atoms(a)%KOCS(k) /= 1) m%kocs(s) = 2
            endif
            if (m%kocs(s) == 2) then
               if (.not. allocated(Target_atoms(a)%Ek)) then
                  why = 'BEB shells need the mean kinetic energies Ek of the shells'
                  return
               endif
               m%Ek(s) = Target_atoms(a)%Ek(k)
            endif
            if (allocated(Target_atoms(a)%KOCS_SHI)) then
               if (Target_atoms(a)%KOCS_SHI(k) /= 1) then
                  why = 'the ion''s BEB cross section (KOCS_SHI /= 1) is not computed by the library'
                  return
               endif
            endif
            m%Ip(s) = Target_atoms(a)%Ip(k)
            m%Nel(s) = Target_atoms(a)%Nel(k)
            m%Auger(s) = Target_atoms(a)%Auger(k)
            m%Radiat(s) = Target_atoms(a)%Radiat(k)
            m%Shl_num(s) = Target_atoms(a)%Shl_num(k)
            m%n_osc(s) = size(Target_atoms(a)%Ritchi(k)%E0)
            do i = 1, m%n_osc(s)
               o = o + 1
               if (o > TRK_MAX_OSC) then
                  why = 'too many CDF oscillators'
                  return
               endif
               m%osc_E0(o) = Target_atoms(a)%Ritchi(k)%E0(i)
               m%osc_A(o) = Target_atoms(a)%Ritchi(k)%A(i)
               m%osc_Gamma(o) = Target_atoms(a)%Ritchi(k)%Gamma(i)
            enddo
         enddo
      enddo
      m%n_phonon = 0
      if (allocated(CDF_Phonon%E0)) then
         m%n_phonon = size(CDF_Phonon%E0)
         if (m%n_phonon > TRK_MAX_PHONON) then
            why = 'too many phonon oscillators'
            return
         endif
         do i = 1, m%n_phonon
            m%ph_E0(i) = CDF_Phonon%E0(i)
            m%ph_A(i) = CDF_Phonon%A(i)
            m%ph_Gamma(i) = CDF_Phonon%Gamma(i)
         enddo
      endif
      m%Dens = Matter%Dens
      m%At_Dens = Matter%At_Dens
      m%Vsound = Matter%Vsound
      m%Layer = Matter%Layer
      m%N_VB_el = Matter%N_VB_el
      m%El_eff_mass = Matter%El_eff_mass
      m%hole_mass = Matter%hole_mass
      m%E_F = Matter%E_F
      m%v_f = Matter%v_f
      m%cut_off = Matter%cut_off
      m%temp = Matter%temp
      m%work_function = Matter%work_function
      m%bar_length = Matter%bar_length
      m%bar_height = Matter%bar_height
      nd = size(Mat_DOS%E)
      m%n_dos = nd
      dos%E = Mat_DOS%E
      dos%DOS = Mat_DOS%DOS
      dos%int_DOS = Mat_DOS%int_DOS
      dos%Eff_m = Mat_DOS%Eff_m
      if (allocated(Mat_DOS%k)) then
         dos%k = Mat_DOS%k
      else
         allocate(dos%k(nd), source=0.0d0)
      endif
      if (allocated(Mat_DOS%k_inv)) then
         dos%k_inv = Mat_DOS%k_inv
      else
         allocate(dos%k_inv(nd), source=0.0d0)
      endif
      if (allocated(Mat_DOS%Eff_m_inv)) then
         dos%Eff_m_inv = Mat_DOS%Eff_m_inv
      else
         allocate(dos%Eff_m_inv(nd), source=0.0d0)
      endif
      m%dos_E = c_loc(dos%E)
      m%dos_DOS = c_loc(dos%DOS)
      m%dos_int_DOS = c_loc(dos%int_DOS)
      m%dos_k = c_loc(dos%k)
      m%dos_Eff_m = c_loc(dos%Eff_m)
      m%dos_k_inv = c_loc(dos%k_inv)
      m%dos_Eff_m_inv = c_loc(dos%Eff_m_inv)
      ok = .true.
   end subroutine flatten_material

   ! SHI (Objects.f90:44-53) into trk_ion; its energy is Basic_particle%E [eV].
   subroutine flatten_ion(SHI, cion)
      type(Ion), intent(in) :: SHI
      type(trk_ion), intent(out) :: cion
      cion%Zat = SHI%Zat
      cion%Mass = SHI%Mass
      cion%E = SHI%E
      cion%Kind_Zeff = SHI%Kind_Zeff
      cion%fixed_Zeff = SHI%fixed_Zeff
      cion%Kind_ion = SHI%Kind_ion
   end subroutine flatten_ion

   ! Flag (Objects.f90:135-180) and the time window into trk_options; the
   ! energy grids end where trk_bridge_e_grid_Emax / _shi_grid_Emax say.
   subroutine flatten_options(NumPar, Tim, dt, o)
      type(Flag), intent(in) :: NumPar
      real(8), intent(in) :: Tim, dt
      type(trk_options), intent(out) :: o
      o%kind_of_EMFP = NumPar%kind_of_EMFP
      o%CDF_elast_Zeff = NumPar%CDF_elast_Zeff
      o%kind_of_DR = NumPar%kind_of_DR
      o%plasmon_Emax = merge(1, 0, NumPar%plasmon_Emax)
      o%include_photons = merge(1, 0, NumPar%include_photons)
      o%dt_flag = NumPar%dt_flag
      o%Tim = Tim
      o%dt = dt
      o%e_grid_Emax = trk_bridge_e_grid_Emax
      o%shi_grid_Emax = trk_bridge_shi_grid_Emax
      o%precision = trk_bridge_precision
      o%batch = 0
      o%block = 0
      o%tally_replicas = 0
      o%gen_capacity = 0
      o%quadrature = 0
      o%reserved = 0
   end subroutine flatten_options

   ! A failed call into the library, as Save_error_details records an error
   ! (Objects.f90:327-350): Err set, the status as Err_Num, the message.
   subroutine report(ctx, status, where, Error_message)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: status
      character(*), intent(in) :: where
      type(Error_handling), intent(inout) :: Error_message
      character(kind=c_char) :: buf(512)
      character(512) :: msg
      integer :: i, rc
      msg = ''
      if (c_associated(ctx)) then
         rc = trk_last_error(ctx, buf, int(512, c_size_t))
         do i = 1, 512
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
         enddo
      endif
      Error_message%Err = .true.
      Error_message%Err_Num = 100 + int(status)
      Error_message%Err_descript = trim(where)//': '//trim(msg)
   end subroutine report

   ! One particle's DSF tables (Differential_MFP, Objects.f90:173-178) into
   ! the library: energies and the E-major rows of dE, dL, dL_emit, dL_absorb.
   ! An empty or ragged set is refused by the library's checks.
   integer(c_int) function pass_dsf(ctx, particle, D) result(rc)
      type(c_ptr), intent(in) :: ctx
      integer, intent(in) :: particle
      type(Differential_MFP), dimension(:), intent(in) :: D
      real(c_double), dimension(:), allocatable :: E, dE, dL, Le, La
      integer :: nE, nW, i
      nE = size(D)
      nW = 0
      if (nE > 0) nW = size(D(1)%dE)
      do i = 1, nE
         if (size(D(i)%dE) /= nW .or. size(D(i)%dL) /= nW .or. size(D(i)%dL_emit) /= nW .or. &
             size(D(i)%dL_absorb) /= nW) nW = 0
      enddo
      allocate(E(max(nE, 1)), dE(max(nE * nW, 1)), dL(max(nE * nW, 1)), Le(max(nE * nW, 1)), La(max(nE * nW, 1)))
      do i = 1, nE
         E(i) = D(i)%E
         if (nW > 0) then
            dE((i - 1) * nW + 1 : i * nW) = D(i)%dE
            dL((i - 1) * nW + 1 : i * nW) = D(i)%dL
            Le((i - 1) * nW + 1 : i * nW) = D(i)%dL_emit
            La((i - 1) * nW + 1 : i * nW) = D(i)%dL_absorb
         endif
      enddo
      rc = trk_dsf_set(ctx, int(particle, c_int32_t), int(nE, c_int32_t), int(nW, c_int32_t), E, dE, dL, Le, La)
   end function pass_dsf

   ! do_Monte_Carlo with the reference's argument list (Monte_Carlo.f90:39-92).
   subroutine do_Monte_Carlo(NMC, SHI, SHI_MFP, diff_SHI_MFP, Target_atoms, Lowest_Ip_At, Lowest_Ip_Shl, CDF_Phonon, &
        Total_el_MFPs, Elastic_MFP, Total_Hole_MFPs, Elastic_Hole_MFP, Total_Photon_MFPs, Mat_DOS, Tim, dt, Matter, &
        NumPar, aidCS, Out_R, Out_V, Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat, &
        Out_nh, Out_Eh, Out_Ehkin, Out_tot_Ne, Out_tot_Nphot, Out_tot_E, &
        Out_E_e, Out_E_phot, Out_E_at, Out_E_h, Out_Eat_dens, Out_theta, Out_theta_h, Out_theta1, Out_Ne_Em, Out_E_Em, &
        Out_Ee_vs_E_Em, Error_message, DSF_DEMFP, DSF_DEMFP_H, Out_field_all, Out_E_field, Out_diff_coeff, MPI_param)
      integer, intent(in) :: NMC
      type(Ion), intent(in) :: SHI
      type(All_MFP), dimension(:), allocatable, intent(in) :: SHI_MFP
      type(All_MFP), dimension(:), allocatable, intent(in) :: diff_SHI_MFP
      type(Atom), dimension(:), intent(in) :: Target_atoms
      integer, intent(in) :: Lowest_Ip_At, Lowest_Ip_Shl
      type(CDF), intent(in) :: CDF_Phonon
      real(8), intent(in) :: Tim
      real(8), intent(in) :: dt
      type(All_MFP), dimension(:), intent(in), target :: Total_el_MFPs
      type(MFP_elastic), intent(in), target :: Elastic_MFP
      type(All_MFP), dimension(:), intent(in), target :: Total_Hole_MFPs
      type(MFP_elastic), intent(in), target :: Elastic_Hole_MFP
      type(All_MFP), dimension(:), intent(in), target :: Total_Photon_MFPs
      type(Solid), intent(in) :: Matter
      type(Density_of_states), intent(in) :: Mat_DOS
      type(Flag), intent(inout) :: NumPar
      real(8), dimension(:), intent(inout), target, contiguous :: Out_R
      real(8), dimension(:,:), intent(inout), target, contiguous :: Out_ne, Out_Ee, Out_nphot, Out_Ephot
      real(8), dimension(:,:), intent(inout), target, contiguous :: Out_Ee_vs_E, Out_Eh_vs_E
      real(8), dimension(:,:), intent(inout), target, contiguous :: Out_Elat
      real(8), dimension(:,:), intent(inout) :: Out_Eat_dens
      real(8), dimension(:,:,:,:), intent(inout), target, contiguous :: Out_nh, Out_Eh, Out_Ehkin
      real(8), dimension(:), intent(inout), target, contiguous :: Out_V
      real(8), dimension(:), intent(inout), target, contiguous :: Out_tot_Ne, Out_tot_Nphot, Out_tot_E
      real(8), dimension(:), intent(inout), target, contiguous :: Out_E_e, Out_E_phot, Out_E_at
      real(8), dimension(:,:,:), intent(inout), target, contiguous :: Out_E_h
      real(8), dimension(:,:), intent(inout), target, contiguous :: Out_theta, Out_theta_h
      real(8), dimension(:), intent(inout) :: Out_theta1
      real(8), dimension(:,:), intent(inout) :: Out_field_all
      real(8), dimension(:,:), intent(inout), target, contiguous :: Out_Ee_vs_E_Em
      real(8), dimension(:), intent(inout), target, contiguous :: Out_Ne_Em
      real(8), dimension(:), intent(inout), target, contiguous :: Out_E_Em
      real(8), dimension(:), intent(inout) :: Out_E_field
      type(Error_handling), intent(inout) :: Error_message
      type(Differential_MFP), dimension(:), intent(in) :: DSF_DEMFP, DSF_DEMFP_H
      type(All_diff_CS), intent(in) :: aidCS
      real(8), dimension(:), intent(inout), target, contiguous :: Out_diff_coeff
      type(Used_MPI_parameters), intent(inout) :: MPI_param
      !------------------------------------------
      type(trk_material) :: m
      type(trk_ion) :: cion
      type(trk_options) :: o
      type(trk_tallies) :: t
      type(bridge_dos), target :: dos
      type(c_ptr) :: ctx
      integer(c_int) :: rc
      integer(c_int32_t) :: Nt
      logical :: ok
      character(200) :: why

      ctx = c_null_ptr
      call flatten_material(Target_atoms, Matter, Mat_DOS, CDF_Phonon, Lowest_Ip_At, Lowest_Ip_Shl, m, dos, ok, why)
      if (.not. ok) then
         Error_message%Err = .true.
         Error_message%Err_Num = 101
         Error_message%Err_descript = 'trekis_gpu_bridge: '//trim(why)
         return
      endif
      call flatten_ion(SHI, cion)
      call flatten_options(NumPar, Tim, dt, o)
      rc = trk_ctx_create(m, cion, o, ctx)
      if (rc /= 0) then
         call report(ctx, rc, 'trk_ctx_create', Error_message)
         call trk_ctx_destroy(ctx)
         return
      endif
      if (NumPar%kind_of_EMFP == 2) then
         ! DSF_DEMFP / DSF_DEMFP_H as reading_DSF_cross_sections left them
         rc = pass_dsf(ctx, 0, DSF_DEMFP)
         if (rc == 0) rc = pass_dsf(ctx, 1, DSF_DEMFP_H)
         if (rc /= 0) then
            call report(ctx, rc, 'trk_dsf_set', Error_message)
            call trk_ctx_destroy(ctx)
            return
         endif
      endif
      rc = trk_build_tables(ctx)
      if (rc /= 0) then
         call report(ctx, rc, 'trk_build_tables', Error_message)
         call trk_ctx_destroy(ctx)
         return
      endif
      rc = trk_time_nodes(ctx, Nt)
      if (size(Out_ne, 1) /= Nt) then
         Error_message%Err = .true.
         Error_message%Err_Num = 101
         Error_message%Err_descript = 'trekis_gpu_bridge: the Out_* arrays do not have the run''s time nodes'
         call trk_ctx_destroy(ctx)
         return
      endif
      t%Nt = Nt
      t%NR = size(Out_R)
      t%n_dos = size(Out_Eh_vs_E, 2)
      t%max_shells = size(Out_nh, 4)
      t%Out_ne = c_loc(Out_ne)
      t%Out_Ee = c_loc(Out_Ee)
      t%Out_nphot = c_loc(Out_nphot)
      t%Out_Ephot = c_loc(Out_Ephot)
      t%Out_Ee_vs_E = c_loc(Out_Ee_vs_E)
      t%Out_Eh_vs_E = c_loc(Out_Eh_vs_E)
      t%Out_Elat = c_loc(Out_Elat)
      t%Out_nh = c_loc(Out_nh)
      t%Out_Eh = c_loc(Out_Eh)
      t%Out_Ehkin = c_loc(Out_Ehkin)
      t%Out_tot_Ne = c_loc(Out_tot_Ne)
      t%Out_tot_Nphot = c_loc(Out_tot_Nphot)
      t%Out_tot_E = c_loc(Out_tot_E)
      t%Out_E_e = c_loc(Out_E_e)
      t%Out_E_phot = c_loc(Out_E_phot)
      t%Out_E_at = c_loc(Out_E_at)
      t%Out_E_h = c_loc(Out_E_h)
      t%Out_theta = c_loc(Out_theta)
      t%Out_theta_h = c_loc(Out_theta_h)
      t%Out_diff_coeff = c_loc(Out_diff_coeff)
      t%Out_Ne_Em = c_loc(Out_Ne_Em)
      t%Out_E_Em = c_loc(Out_E_Em)
      t%Out_Ee_vs_E_Em = c_loc(Out_Ee_vs_E_Em)
      t%n_events = 0
      t%n_particles = 0
      rc = trk_run_mc(ctx, int(NMC, c_int64_t), trk_bridge_seed, trk_bridge_first_iter, Out_R, Out_V, t%NR, t)
      if (rc /= 0) call report(ctx, rc, 'trk_run_mc', Error_message)
      call trk_ctx_destroy(ctx)
   end subroutine do_Monte_Carlo

end module trekis_gpu_bridge
