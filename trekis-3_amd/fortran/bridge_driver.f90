! bridge_driver.f90 -- what the reference's MAIN does around do_Monte_Carlo
! (Universal_MC_for_SHI_MAIN.f90:128-276), reduced to the bridge: it fills the
! reference's own types (Objects.f90) from a descriptor file, allocates the
! Out_* arrays with the dimensions of Allocate_out_arrays
! (Sorting_output_data.f90:1144-1250), calls do_Monte_Carlo from
! trekis_gpu_bridge and writes the arrays out.
!
! Usage: bridge_driver <descriptor file> <output file>
! The descriptor file is a stream of native integers (int32, int64) and
! doubles in the order read below; the output file holds the raw doubles of
! Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat,
! Out_nh, Out_Eh, Out_Ehkin, Out_tot_Ne, Out_tot_Nphot, Out_tot_E, Out_E_e,
! Out_E_phot, Out_E_at, Out_E_h, Out_theta, Out_theta_h, Out_diff_coeff,
! Out_Ne_Em, Out_E_Em, Out_Ee_vs_E_Em in Fortran order, after the error flag
! and number (two int32).  With DSF elastic scattering (kind_of_EMFP = 2) the
! descriptor ends with the electron and hole DSF tables (read_dsf).
program bridge_driver
   use iso_c_binding
   use Objects
   use trekis_gpu_bridge
   implicit none
   type(Atom), dimension(:), allocatable :: Target_atoms
   type(Solid) :: Matter
   type(Density_of_states) :: Mat_DOS
   type(CDF) :: CDF_Phonon
   type(Ion) :: SHI
   type(Flag) :: NumPar
   type(Error_handling) :: Error_message
   type(Used_MPI_parameters) :: MPI_param
   type(All_MFP), dimension(:), allocatable :: SHI_MFP, diff_SHI_MFP, Total_el_MFPs, Total_Hole_MFPs, Total_Photon_MFPs
   type(MFP_elastic) :: Elastic_MFP, Elastic_Hole_MFP
   type(All_diff_CS) :: aidCS
   type(Differential_MFP), dimension(:), allocatable :: DSF_DEMFP, DSF_DEMFP_H
   real(8), dimension(:), allocatable :: Out_R, Out_V, Out_tot_Ne, Out_tot_Nphot, Out_tot_E, Out_E_e, Out_E_phot
   real(8), dimension(:), allocatable :: Out_E_at, Out_theta1, Out_Ne_Em, Out_E_Em, Out_E_field, Out_diff_coeff
   real(8), dimension(:,:), allocatable :: Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat
   real(8), dimension(:,:), allocatable :: Out_Eat_dens, Out_theta, Out_theta_h, Out_field_all, Out_Ee_vs_E_Em
   real(8), dimension(:,:,:), allocatable :: Out_E_h
   real(8), dimension(:,:,:,:), allocatable :: Out_nh, Out_Eh, Out_Ehkin
   character(512) :: fin, fout
   integer(c_int32_t) :: na, nsh, nosc, nph, nd, i4, Nt, NR, ms, NMC, ipl, iph
   integer(c_int64_t) :: seed
   real(c_double) :: Tim, dt, eg, sg
   integer :: a, k, u
   real(8) :: r8(3)

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
   read(u) na
   allocate(Target_atoms(na))
   do a = 1, na
      read(u) Target_atoms(a)%Zat, Target_atoms(a)%Mass, Target_atoms(a)%Pers, nsh
      Target_atoms(a)%N_shl = nsh
      allocate(Target_atoms(a)%Ip(nsh), Target_atoms(a)%Nel(nsh), Target_atoms(a)%Auger(nsh))
      allocate(Target_atoms(a)%Radiat(nsh), Target_atoms(a)%Shl_num(nsh), Target_atoms(a)%Ritchi(nsh))
      allocate(Target_atoms(a)%KOCS(nsh), Target_atoms(a)%Ek(nsh))
      Target_atoms(a)%KOCS = 1
      do k = 1, nsh
         read(u) Target_atoms(a)%Ip(k), Target_atoms(a)%Nel(k), Target_atoms(a)%Auger(k), Target_atoms(a)%Radiat(k)
         read(u) i4, nosc, Target_atoms(a)%KOCS(k)
         read(u) Target_atoms(a)%Ek(k)
         Target_atoms(a)%Shl_num(k) = i4
         allocate(Target_atoms(a)%Ritchi(k)%E0(nosc), Target_atoms(a)%Ritchi(k)%A(nosc))
         allocate(Target_atoms(a)%Ritchi(k)%Gamma(nosc))
         do i4 = 1, nosc
            read(u) r8
            Target_atoms(a)%Ritchi(k)%E0(i4) = r8(1)
            Target_atoms(a)%Ritchi(k)%A(i4) = r8(2)
            Target_atoms(a)%Ritchi(k)%Gamma(i4) = r8(3)
         enddo
      enddo
   enddo
   read(u) nph
   if (nph > 0) then
      allocate(CDF_Phonon%E0(nph), CDF_Phonon%A(nph), CDF_Phonon%Gamma(nph))
      do i4 = 1, nph
         read(u) r8
         CDF_Phonon%E0(i4) = r8(1)
         CDF_Phonon%A(i4) = r8(2)
         CDF_Phonon%Gamma(i4) = r8(3)
      enddo
   endif
   read(u) Matter%Dens, Matter%At_Dens, Matter%Vsound, Matter%Layer, Matter%N_VB_el, Matter%El_eff_mass, &
           Matter%hole_mass, Matter%E_F, Matter%v_f, Matter%cut_off, Matter%temp, Matter%work_function, &
           Matter%bar_length, Matter%bar_height
   read(u) nd
   allocate(Mat_DOS%E(nd), Mat_DOS%DOS(nd), Mat_DOS%int_DOS(nd), Mat_DOS%k(nd), Mat_DOS%Eff_m(nd))
   allocate(Mat_DOS%k_inv(nd), Mat_DOS%Eff_m_inv(nd))
   read(u) Mat_DOS%E, Mat_DOS%DOS, Mat_DOS%int_DOS, Mat_DOS%k, Mat_DOS%Eff_m, Mat_DOS%k_inv, Mat_DOS%Eff_m_inv
   read(u) SHI%Zat, SHI%Mass, SHI%E, SHI%Kind_Zeff, SHI%fixed_Zeff, SHI%Kind_ion
   read(u) NumPar%kind_of_EMFP, NumPar%CDF_elast_Zeff, NumPar%kind_of_DR, ipl, iph, NumPar%dt_flag
   NumPar%plasmon_Emax = (ipl /= 0)
   NumPar%include_photons = (iph /= 0)
   read(u) Tim, dt, eg, sg
   read(u) NMC, seed, Nt, NR, ms
   allocate(Out_R(NR), Out_V(NR))
   read(u) Out_R, Out_V
   if (NumPar%kind_of_EMFP == 2) then
      ! DSF_DEMFP and DSF_DEMFP_H as reading_DSF_cross_sections leaves them
      call read_dsf(u, DSF_DEMFP)
      call read_dsf(u, DSF_DEMFP_H)
   else
      allocate(DSF_DEMFP(0), DSF_DEMFP_H(0))
   endif
   close(u)
   trk_bridge_seed = seed
   trk_bridge_e_grid_Emax = eg
   trk_bridge_shi_grid_Emax = sg

   ! Allocate_out_arrays, Sorting_output_data.f90:1191-1249
   allocate(Out_ne(Nt, NR), Out_Ee(Nt, NR), Out_nphot(Nt, NR), Out_Ephot(Nt, NR), Out_Elat(Nt, NR))
   allocate(Out_Ee_vs_E(Nt, NR), Out_Eh_vs_E(Nt, nd), Out_Eat_dens(Nt, NR), Out_field_all(Nt, NR))
   allocate(Out_Ee_vs_E_Em(Nt, NR))
   allocate(Out_nh(Nt, NR, na, ms), Out_Eh(Nt, NR, na, ms), Out_Ehkin(Nt, NR, na, ms), Out_E_h(Nt, na, ms))
   allocate(Out_tot_Ne(Nt), Out_tot_Nphot(Nt), Out_tot_E(Nt), Out_E_e(Nt), Out_E_phot(Nt), Out_E_at(Nt))
   allocate(Out_Ne_Em(Nt), Out_E_Em(Nt), Out_E_field(Nt), Out_diff_coeff(Nt), Out_theta1(Nt))
   allocate(Out_theta(Nt + 1, 180), Out_theta_h(Nt + 1, 180))
   Out_ne = 0.0d0; Out_Ee = 0.0d0; Out_nphot = 0.0d0; Out_Ephot = 0.0d0; Out_Elat = 0.0d0
   Out_Ee_vs_E = 0.0d0; Out_Eh_vs_E = 0.0d0; Out_Eat_dens = 0.0d0; Out_field_all = 0.0d0; Out_Ee_vs_E_Em = 0.0d0
   Out_nh = 0.0d0; Out_Eh = 0.0d0; Out_Ehkin = 0.0d0; Out_E_h = 0.0d0
   Out_tot_Ne = 0.0d0; Out_tot_Nphot = 0.0d0; Out_tot_E = 0.0d0; Out_E_e = 0.0d0; Out_E_phot = 0.0d0
   Out_E_at = 0.0d0; Out_Ne_Em = 0.0d0; Out_E_Em = 0.0d0; Out_E_field = 0.0d0; Out_diff_coeff = 0.0d0
   Out_theta1 = 0.0d0; Out_theta = 0.0d0; Out_theta_h = 0.0d0
   Error_message%Err = .false.
   Error_message%Err_Num = 0
   allocate(SHI_MFP(0), diff_SHI_MFP(0), Total_el_MFPs(0), Total_Hole_MFPs(0), Total_Photon_MFPs(0))

   ! Universal_MC_for_SHI_MAIN.f90:271-276
   call do_Monte_Carlo(NMC, SHI, SHI_MFP, diff_SHI_MFP, Target_atoms, 1, Target_atoms(1)%N_shl, CDF_Phonon, &
        Total_el_MFPs, Elastic_MFP, Total_Hole_MFPs, Elastic_Hole_MFP, Total_Photon_MFPs, Mat_DOS, Tim, dt, Matter, &
        NumPar, aidCS, Out_R, Out_V, Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat, &
        Out_nh, Out_Eh, Out_Ehkin, Out_tot_Ne, Out_tot_Nphot, Out_tot_E, &
        Out_E_e, Out_E_phot, Out_E_at, Out_E_h, Out_Eat_dens, Out_theta, Out_theta_h, Out_theta1, Out_Ne_Em, Out_E_Em, &
        Out_Ee_vs_E_Em, Error_message, DSF_DEMFP, DSF_DEMFP_H, Out_field_all, Out_E_field, Out_diff_coeff, MPI_param)
   if (Error_message%Err) print '(a)', trim(Error_message%Err_descript)

   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   write(u) merge(1_c_int32_t, 0_c_int32_t, Error_message%Err), int(Error_message%Err_Num, c_int32_t)
   write(u) Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat, Out_nh, Out_Eh, Out_Ehkin
   write(u) Out_tot_Ne, Out_tot_Nphot, Out_tot_E, Out_E_e, Out_E_phot, Out_E_at, Out_E_h, Out_theta, Out_theta_h
   write(u) Out_diff_coeff
   write(u) Out_Ne_Em, Out_E_Em, Out_Ee_vs_E_Em
   close(u)

contains

   ! one particle's DSF_DEMFP: nE, nW (int32), then per energy E and the
   ! rows dE, dL, dL_emit, dL_absorb (nW doubles each)
   subroutine read_dsf(u, D)
      integer, intent(in) :: u
      type(Differential_MFP), dimension(:), allocatable, intent(inout) :: D
      integer(c_int32_t) :: nE, nW
      integer :: i
      read(u) nE, nW
      allocate(D(nE))
      do i = 1, nE
         allocate(D(i)%dE(nW), D(i)%dL(nW), D(i)%dL_emit(nW), D(i)%dL_absorb(nW))
         read(u) D(i)%E, D(i)%dE, D(i)%dL, D(i)%dL_emit, D(i)%dL_absorb
      enddo
   end subroutine read_dsf
end program bridge_driver
