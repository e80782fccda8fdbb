! trekis_gpu_c.f90 -- ISO_C_BINDING interface to libtrekis_gpu.so.
!
! Mirrors include/trekis_gpu.h type for type.  A Fortran driver (the
! reference's MAIN, Universal_MC_for_SHI_MAIN.f90) uses this module from a
! bridge that flattens its Atom / Solid / Density_of_states / Flag objects
! (Objects.f90:44-312) into these types and calls trk_run_mc in place of
! do_Monte_Carlo (Monte_Carlo.f90:39-92).
module trekis_gpu_c
   use iso_c_binding
   implicit none

   integer, parameter :: TRK_MAX_ATOMS = 8, TRK_MAX_SHELLS = 32, TRK_MAX_OSC = 128, TRK_MAX_PHONON = 16

   type, bind(C) :: trk_material
      integer(c_int32_t) :: n_atoms
      integer(c_int32_t) :: Zat(TRK_MAX_ATOMS)
      real(c_double)     :: Mass(TRK_MAX_ATOMS)
      real(c_double)     :: Pers(TRK_MAX_ATOMS)
      integer(c_int32_t) :: n_shells(TRK_MAX_ATOMS)
      real(c_double)     :: Ip(TRK_MAX_SHELLS)
      real(c_double)     :: Nel(TRK_MAX_SHELLS)
      real(c_double)     :: Auger(TRK_MAX_SHELLS)
      real(c_double)     :: Radiat(TRK_MAX_SHELLS)
      integer(c_int32_t) :: Shl_num(TRK_MAX_SHELLS)
      integer(c_int32_t) :: n_osc(TRK_MAX_SHELLS)
      real(c_double)     :: osc_E0(TRK_MAX_OSC)
      real(c_double)     :: osc_A(TRK_MAX_OSC)
      real(c_double)     :: osc_Gamma(TRK_MAX_OSC)
      integer(c_int32_t) :: n_phonon
      real(c_double)     :: ph_E0(TRK_MAX_PHONON)
      real(c_double)     :: ph_A(TRK_MAX_PHONON)
      real(c_double)     :: ph_Gamma(TRK_MAX_PHONON)
      real(c_double)     :: Dens, At_Dens, Vsound, Layer, N_VB_el, El_eff_mass, hole_mass
      real(c_double)     :: E_F, v_f, cut_off, temp, work_function, bar_length, bar_height
      integer(c_int32_t) :: n_dos
      type(c_ptr)        :: dos_E, dos_DOS, dos_int_DOS, dos_k, dos_Eff_m, dos_k_inv, dos_Eff_m_inv
      integer(c_int32_t) :: kocs(TRK_MAX_SHELLS)
      real(c_double)     :: Ek(TRK_MAX_SHELLS)
   end type trk_material

   type, bind(C) :: trk_ion
      integer(c_int32_t) :: Zat
      real(c_double)     :: Mass
      real(c_double)     :: E
      integer(c_int32_t) :: Kind_Zeff
      real(c_double)     :: fixed_Zeff
      integer(c_int32_t) :: Kind_ion
   end type trk_ion

   type, bind(C) :: trk_options
      integer(c_int32_t) :: kind_of_EMFP, CDF_elast_Zeff, kind_of_DR, plasmon_Emax, include_photons, dt_flag
      real(c_double)     :: Tim, dt, e_grid_Emax, shi_grid_Emax
      integer(c_int32_t) :: precision, batch, block, tally_replicas, gen_capacity
      integer(c_int32_t) :: quadrature
      integer(c_int32_t) :: reserved(2)
   end type trk_options

   type, bind(C) :: trk_tallies
      integer(c_int32_t) :: Nt, NR, n_dos, max_shells
      type(c_ptr) :: Out_ne, Out_Ee, Out_nphot, Out_Ephot, Out_Ee_vs_E, Out_Eh_vs_E, Out_Elat
      type(c_ptr) :: Out_nh, Out_Eh, Out_Ehkin, Out_tot_Ne, Out_tot_Nphot, Out_tot_E
      type(c_ptr) :: Out_E_e, Out_E_phot, Out_E_at, Out_E_h, Out_theta, Out_theta_h, Out_diff_coeff
      type(c_ptr) :: Out_Ne_Em, Out_E_Em, Out_Ee_vs_E_Em
      integer(c_int64_t) :: n_events(8)
      integer(c_int64_t) :: n_particles
   end type trk_tallies

   interface
      integer(c_int) function trk_ctx_create(mat, ion, opt, ctx) bind(C, name="trk_ctx_create")
         import :: c_int, c_ptr, trk_material, trk_ion, trk_options
         type(trk_material), intent(in) :: mat
         type(trk_ion), intent(in) :: ion
         type(trk_options), intent(in) :: opt
         type(c_ptr), intent(out) :: ctx
      end function
      subroutine trk_ctx_destroy(ctx) bind(C, name="trk_ctx_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine
      integer(c_int) function trk_last_error(ctx, buf, n) bind(C, name="trk_last_error")
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_size_t), value :: n
      end function
      integer(c_int) function trk_dsf_set(ctx, particle, nE, nW, E, dE, dL, dL_emit, dL_absorb) &
            bind(C, name="trk_dsf_set")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: particle, nE, nW
         real(c_double), intent(in) :: E(*), dE(*), dL(*), dL_emit(*), dL_absorb(*)
      end function
      integer(c_int) function trk_build_tables(ctx) bind(C, name="trk_build_tables")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function trk_time_nodes(ctx, Nt) bind(C, name="trk_time_nodes")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: Nt
      end function
      integer(c_int) function trk_run_mc(ctx, nmc, seed, first_iter, Out_R, Out_V, NR, tallies) &
            bind(C, name="trk_run_mc")
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double, trk_tallies
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: nmc, seed, first_iter
         real(c_double), intent(in) :: Out_R(*), Out_V(*)
         integer(c_int32_t), value :: NR
         type(trk_tallies), intent(inout) :: tallies
      end function
      integer(c_int) function trk_sumrules(A, E0, G, n, x_min, Omega, ksum, fsum) bind(C, name="trk_sumrules")
         import :: c_int, c_double, c_int32_t
         real(c_double), intent(in) :: A(*), E0(*), G(*)
         integer(c_int32_t), value :: n
         real(c_double), value :: x_min, Omega
         real(c_double), intent(out) :: ksum, fsum
      end function
      real(c_double) function trk_rng_uniform(seed, id, iter, n) bind(C, name="trk_rng_uniform")
         import :: c_double, c_int64_t, c_int32_t
         integer(c_int64_t), value :: seed, id
         integer(c_int32_t), value :: iter, n
      end function
   end interface
end module trekis_gpu_c
