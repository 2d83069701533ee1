! shcfd_f.f90 -- Fortran host binding of libshcfd (include/shcfd.h).
!
! ISO_C_BINDING interfaces to the C ABI, plus wrappers that carry the
! reference's subroutine names (set_w, set_thermo_prop, set_BC, set_HV,
! set_TG, set_TGv, set_dt, set_ABpm, set_dqdt, calc_next_step_implicit,
! calc_next_step, calc_next_step_RK2_first/second, check_convergence), so a
! main loop written like the reference's generated main program
! (store/core/main.body.raw.f90, store/time_scheme/*/main.main.raw.f90)
! drives the GPU kernels unchanged.  As in the reference, state is a module
! global (here the device context handle); a nonzero status prints the
! reference's message and stops.
module shcfd_f
   use iso_c_binding
   implicit none

   type, bind(C) :: shcfd_params_t
      integer(c_int) :: scheme, time_step, high_order, limiter, viscous, thermal
      integer(c_int) :: num_internal_loop, reaction
      real(c_double) :: phi_mus, k_mus, cfl, cfl_tau, omega_max, omega_min, dqmax
      real(c_double) :: res_rate_warning, res_rate_error, kappa, rgas, nu
      integer(c_int) :: dimension   ! 0 two-dimensional, 1 axi-symmetric (checkout.inp "dimension")
   end type shcfd_params_t

   ! field identifiers (shcfd_types.h)
   integer(c_int), parameter :: SHCFD_F_Q = 0, SHCFD_F_QP = 1, SHCFD_F_QPP = 2, SHCFD_F_W = 3
   integer(c_int), parameter :: SHCFD_F_XH = 27, SHCFD_F_RH = 28
   integer(c_int), parameter :: SCHEME_EULER = 0, SCHEME_RK2 = 1, SCHEME_LUSGS = 2
   integer(c_int), parameter :: SCHEME_NR = 3, SCHEME_PRECON = 4
   integer(c_int), parameter :: SCHEME_DUAL = 5, SCHEME_PRECON_LUSGS = 6

   interface
      integer(c_int) function shcfd_create(nblocks, ni, nj, ny, params, out) bind(C, name="shcfd_create")
         import :: c_int, c_ptr, shcfd_params_t
         integer(c_int), value :: nblocks, ny
         integer(c_int), intent(in) :: ni(*), nj(*)
         type(shcfd_params_t), intent(in) :: params
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function shcfd_destroy(ctx) bind(C, name="shcfd_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_upload_field(ctx, field, block, host) bind(C, name="shcfd_upload_field")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field, block
         real(c_double), intent(in) :: host(*)
      end function
      integer(c_int) function shcfd_download_field(ctx, field, block, host) bind(C, name="shcfd_download_field")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field, block
         real(c_double), intent(out) :: host(*)
      end function
      integer(c_int) function shcfd_set_cut_table(ctx, rows, nrows) bind(C, name="shcfd_set_cut_table")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(in) :: rows(*)
         integer(c_int), value :: nrows
      end function
      integer(c_int) function shcfd_set_bc_table(ctx, rows, nrows, pool, npool) bind(C, name="shcfd_set_bc_table")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), intent(in) :: rows(*)
         integer(c_int), value :: nrows, npool
         real(c_double), intent(in) :: pool(*)
      end function
      integer(c_int) function shcfd_init_geometry(ctx) bind(C, name="shcfd_init_geometry")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_w_c(ctx) bind(C, name="shcfd_set_w")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_thermo_prop_c(ctx, step) bind(C, name="shcfd_set_thermo_prop")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: step
      end function
      integer(c_int) function shcfd_set_bc_c(ctx, step) bind(C, name="shcfd_set_bc")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: step
      end function
      integer(c_int) function shcfd_set_hv_c(ctx) bind(C, name="shcfd_set_hv")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_tg_c(ctx) bind(C, name="shcfd_set_tg")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_tgv_c(ctx) bind(C, name="shcfd_set_tgv")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_dt_c(ctx, step, dt) bind(C, name="shcfd_set_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: step
         real(c_double), intent(out) :: dt
      end function
      integer(c_int) function shcfd_set_abpm_c(ctx) bind(C, name="shcfd_set_abpm")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_dqdt_c(ctx) bind(C, name="shcfd_set_dqdt")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_calc_next_step_implicit_c(ctx, step_internal, res, omega) &
            bind(C, name="shcfd_calc_next_step_implicit")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: step_internal
         real(c_double), intent(out) :: res, omega
      end function
      integer(c_int) function shcfd_calc_next_step_c(ctx) bind(C, name="shcfd_calc_next_step")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_calc_next_step_rk2_c(ctx, stage) bind(C, name="shcfd_calc_next_step_rk2")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: stage
      end function
      integer(c_int) function shcfd_copy_state(ctx, dst, src) bind(C, name="shcfd_copy_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: dst, src
      end function
      integer(c_int) function shcfd_energy_rms(ctx, rms) bind(C, name="shcfd_energy_rms")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: rms
      end function
      integer(c_int) function shcfd_set_time(ctx, tt, first_dual) bind(C, name="shcfd_set_time")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: tt
         integer(c_int), value :: first_dual
      end function
      integer(c_int) function shcfd_get_time(ctx, dt, tt) bind(C, name="shcfd_get_time")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: dt, tt
      end function
      integer(c_int) function shcfd_load_mechanism(ctx, chem_inp, mw_inp, trans_inp, rtol, atol) &
            bind(C, name="shcfd_load_mechanism")
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: chem_inp(*), mw_inp(*), trans_inp(*)
         real(c_double), value :: rtol, atol
      end function
      integer(c_int) function shcfd_load_nasa9(ctx, chem_inp, thermo_inp, trans_inp, control_inp) &
            bind(C, name="shcfd_load_nasa9")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: chem_inp(*), thermo_inp(*), trans_inp(*), control_inp(*)
      end function
      integer(c_int) function shcfd_proceed_reaction_c(ctx) bind(C, name="shcfd_proceed_reaction")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_check_error(ctx) bind(C, name="shcfd_check_error")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function shcfd_set_option(ctx, option, value) bind(C, name="shcfd_set_option")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: option, value
      end function
   end interface

   ! module state, as the reference's module variable / var_lusgs
   type(c_ptr), save :: ctx = c_null_ptr
   type(shcfd_params_t), save :: prm
   integer, save :: nplane = 0, ny_ = 1
   double precision, save :: res, res1, omega
   logical, save :: first_dual = .true.
   double precision, save :: max_step = 0d0, min_RMS = 0d0
   integer, save :: OutPeriod = 1
   double precision, save :: dt_grbl = 0d0, tt = 0d0

contains

   subroutine shcfd_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      if (rc /= 0) then
         select case (rc)
         case (5)
            print *, "Error: Omega becomes too small. (", where, ")"
         case (4)
            print *, "res is NaN (", where, ")"
         case (6)
            print *, "ERROR. EXCEED MAXIMUM RESIDUAL."
         case (7)
            print '(a)', "Error: Exceed max_T_loop at temperature determination iteration"
         case (8)
            print *, "not converted at CHEMEQ2 (", where, ")"
         case default
            print *, "shcfd error ", rc, " in ", where
         end select
         stop 1
      end if
   end subroutine shcfd_check

   ! --- reference-named drop-ins -------------------------------------------
   subroutine set_w()                       ! store/core/scheme.f90:1-29
      call shcfd_check(shcfd_set_w_c(ctx), "set_w")
   end subroutine
   subroutine set_thermo_prop(step)         ! therm_lib/ideal/thermal_model.raw.f90:7-33,
      integer, intent(in) :: step           ! therm_lib/chemkin/flow/reactive/thermal_model.f90:13-42
      call shcfd_check(shcfd_set_thermo_prop_c(ctx, int(step, c_int)), "set_thermo_prop")
   end subroutine
   subroutine proceed_reaction()            ! therm_lib/chemkin/flow/reactive/thermal_model.f90:45-73
      call shcfd_check(shcfd_proceed_reaction_c(ctx), "proceed_reaction")
      call shcfd_check(shcfd_check_error(ctx), "proceed_reaction")
   end subroutine
   subroutine init_therm(chem_inp, mw_inp, trans_inp, rtol, atol)   ! reactive main.part_init.f90
      character(*), intent(in) :: chem_inp, mw_inp, trans_inp
      double precision, intent(in) :: rtol, atol
      call shcfd_check(shcfd_load_mechanism(ctx, trim(chem_inp)//c_null_char, trim(mw_inp)//c_null_char, &
                                            trim(trans_inp)//c_null_char, rtol, atol), "init_therm")
   end subroutine
   subroutine init_pack_cold_flow(chem_inp, thermo_inp, trans_inp, control_inp)  ! NASA/core/sub_chem.f90:380-387
      character(*), intent(in) :: chem_inp, thermo_inp, trans_inp, control_inp
      call shcfd_check(shcfd_load_nasa9(ctx, trim(chem_inp)//c_null_char, trim(thermo_inp)//c_null_char, &
                                        trim(trans_inp)//c_null_char, trim(control_inp)//c_null_char), &
                       "init_pack_cold_flow")
   end subroutine
   subroutine set_BC(step)                  ! store/cond/no-nV/condition.tail.f90
      integer, intent(in) :: step
      call shcfd_check(shcfd_set_bc_c(ctx, int(step, c_int)), "set_BC")
   end subroutine
   subroutine set_HV()                      ! store/high_order/muscl/muscl.f90:7-73
      call shcfd_check(shcfd_set_hv_c(ctx), "set_HV")
   end subroutine
   subroutine set_TG()                      ! store/core/scheme.f90:31-259
      call shcfd_check(shcfd_set_tg_c(ctx), "set_TG")
   end subroutine
   subroutine set_TGv()                     ! store/viscosity/viscous/no-nV/2d/sch_viscous.f90:56-272
      call shcfd_check(shcfd_set_tgv_c(ctx), "set_TGv")
   end subroutine
   subroutine set_dt(step)                  ! store/time_step/{global,local}/set_dt.f90
      integer, intent(in) :: step
      real(c_double) :: d
      call shcfd_check(shcfd_set_dt_c(ctx, int(step, c_int), d), "set_dt")
      if (prm%time_step == 0) then
         dt_grbl = d
         tt = tt + d
      end if
   end subroutine
   subroutine set_ABpm()                    ! store/time_scheme/*/2d/sch_*.raw.f90 set_ABpm
      call shcfd_check(shcfd_set_abpm_c(ctx), "set_ABpm")
   end subroutine
   subroutine set_dqdt()                    ! store/time_scheme/dual/2d/sch_dual.raw.f90:211-240
      call shcfd_check(shcfd_set_dqdt_c(ctx), "set_dqdt")
   end subroutine
   subroutine calc_next_step_implicit(step_internal)   ! time_*_viscous.f90
      integer, intent(in) :: step_internal
      real(c_double) :: r, o
      call shcfd_check(shcfd_calc_next_step_implicit_c(ctx, int(step_internal, c_int), r, o), &
                       "calc_next_step_implicit")
      res = r
      omega = o
      if (step_internal == 1) res1 = res
   end subroutine
   subroutine calc_next_step()              ! store/time_scheme/euler/time_euler_viscous.raw.f90
      call shcfd_check(shcfd_calc_next_step_c(ctx), "calc_next_step")
   end subroutine
   subroutine calc_next_step_RK2_first()    ! store/time_scheme/RK2/time_RK2_viscous.raw.f90:1-24
      call shcfd_check(shcfd_calc_next_step_rk2_c(ctx, 1_c_int), "calc_next_step_RK2_first")
   end subroutine
   subroutine calc_next_step_RK2_second()   ! store/time_scheme/RK2/time_RK2_viscous.raw.f90:26-59
      call shcfd_check(shcfd_calc_next_step_rk2_c(ctx, 2_c_int), "calc_next_step_RK2_second")
   end subroutine
   subroutine copy_qp()                     ! qp = q
      call shcfd_check(shcfd_copy_state(ctx, SHCFD_F_QP, SHCFD_F_Q), "qp=q")
   end subroutine
   subroutine copy_qpp()                    ! qpp = qp
      call shcfd_check(shcfd_copy_state(ctx, SHCFD_F_QPP, SHCFD_F_QP), "qpp=qp")
   end subroutine
   subroutine check_internal_loop_convergence(step, step_res)   ! sch_dual.raw.f90:284-299
      integer, intent(in) :: step, step_res
      if (res/res1 > prm%res_rate_warning) then
         print '(a,i12,es11.3)', "WARNING:NOT CONVERGED AT INTERNAL LOOP (step,res_rate) = ", step, res/res1
         if (res/res1 > prm%res_rate_error .and. step /= step_res + 1) stop "ERROR. EXCEED MAXIMUM RESIDUAL."
      end if
      first_dual = .false.
      call shcfd_check(shcfd_set_time(ctx, tt, 0_c_int), "set_time")
   end subroutine

   ! store/core/check_convergence.f90:1-59 (RMS of the energy change every
   ! OutPeriod steps, NaN abort, stop on min_RMS after step 5000 or max_step)
   logical function check_convergence(step, step_res)
      integer, intent(in) :: step, step_res
      real(c_double) :: rms
      check_convergence = .true.
      if (mod(step, OutPeriod) /= 0) return
      call shcfd_check(shcfd_energy_rms(ctx, rms), "check_convergence")
      if (rms /= rms) then
         print *, "NaN at procid=", 0
         stop 1
      end if
      open(82, file="RMS.dat", position='append')
      write(82, *) step, rms
      close(82)
      print '(a,i12,x,2(a,es15.7,x))', "step=", step, "log(RMS)=", log10(max(rms, tiny(rms))), "dt_grbl=", dt_grbl
      if (min_RMS >= rms .and. step > 5000) then
         print *, "RMS accomplished at step=", step, " at tmp=", rms
         check_convergence = .false.
      else if (max_step + step_res <= step) then
         print *, "Error:Exceed Max Step. NOT CONVERTED."
         check_convergence = .false.
      end if
   end function check_convergence

end module shcfd_f
