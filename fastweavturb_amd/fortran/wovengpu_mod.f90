! ISO_C_BINDING interface to libwovengpu.so (include/wovengpu.h).
! Scalars are passed by value, arrays by reference; every function returns 0 on success or a
! negative WT_E* code, with the message available from wt_last_error_str.
module wovengpu
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: WT_FIELD_WX = 0, WT_FIELD_WY = 1, WT_FIELD_WZ = 2
   integer(c_int), parameter :: WT_FIELD_UX = 3, WT_FIELD_UY = 4, WT_FIELD_UZ = 5

   ! wt_params: the scalar header of parameters_input.dat (tube.f90:191-216)
   type, bind(C) :: wt_params
      integer(c_int32_t) :: N, n_d, line_s, line_e
      real(c_double) :: sigma_ratio, eta_t
      integer(c_int32_t) :: ntf_times, ntf_pre
      real(c_double) :: Rtube_ratio, dzeta_sup
   end type

   type, bind(C) :: wt_seg_stats
      integer(c_int64_t) :: npoint_all, nsec_all, nsub_all, nseg_all
      integer(c_int32_t) :: ntf_max, reserved
      real(c_double) :: dtf_min
      integer(c_int64_t) :: ntf_sum, theta_exact_sections
   end type

   interface
      integer(c_int) function wt_create(ctx, n, precision, device) bind(C, name="wt_create")
         import :: c_ptr, c_int, c_int32_t
         type(c_ptr), intent(out) :: ctx
         integer(c_int32_t), value :: n, precision, device
      end function
      subroutine wt_destroy(ctx) bind(C, name="wt_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine
      type(c_ptr) function wt_last_error(ctx) bind(C, name="wt_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function wt_set_params(ctx, p) bind(C, name="wt_set_params")
         import :: c_ptr, c_int, wt_params
         type(c_ptr), value :: ctx
         type(wt_params), intent(in) :: p
      end function
      integer(c_int) function wt_set_tubes(ctx, nline, npoint, sigma, gamma, n_T, if_filter, cx, cy, cz) &
         bind(C, name="wt_set_tubes")
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nline
         integer(c_int32_t), intent(in) :: npoint(*), n_T(*), if_filter(*)
         real(c_double), intent(in) :: sigma(*), gamma(*), cx(*), cy(*), cz(*)
      end function
      integer(c_int) function wt_build_segments(ctx, st) bind(C, name="wt_build_segments")
         import :: c_ptr, c_int, wt_seg_stats
         type(c_ptr), value :: ctx
         type(wt_seg_stats), intent(out) :: st
      end function
      integer(c_int) function wt_bin(ctx) bind(C, name="wt_bin")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function wt_deposit(ctx, st) bind(C, name="wt_deposit")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         type(c_ptr), value :: st
      end function
      integer(c_int) function wt_biot_savart(ctx, etot, up, enstrophy) bind(C, name="wt_biot_savart")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: etot, up, enstrophy
      end function
      integer(c_int) function wt_spectrum(ctx, field, E, cap) bind(C, name="wt_spectrum")
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: field, cap
         real(c_double), intent(out) :: E(*)
      end function
      integer(c_int32_t) function wt_nek(n) bind(C, name="wt_nek")
         import :: c_int32_t
         integer(c_int32_t), value :: n
      end function
      integer(c_int32_t) function wt_sp_separations(n, r, cap) bind(C, name="wt_sp_separations")
         import :: c_int32_t
         integer(c_int32_t), value :: n, cap
         integer(c_int32_t), intent(out) :: r(*)
      end function
      integer(c_int) function wt_structure_functions(ctx, mode, pmin, pmax, nsep, seps, S) &
         bind(C, name="wt_structure_functions")
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: mode, pmin, pmax, nsep
         integer(c_int32_t), intent(in) :: seps(*)
         real(c_double), intent(out) :: S(*)
      end function
      integer(c_int) function wt_pdf(ctx, field, nbins, val, pdf, counts) bind(C, name="wt_pdf")
         import :: c_ptr, c_int, c_int32_t, c_float
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: field, nbins
         real(c_float), intent(out) :: val(*), pdf(*)
         type(c_ptr), value :: counts
      end function
      integer(c_int) function wt_get_field(ctx, which, host, host_is_fp64) bind(C, name="wt_get_field")
         import :: c_ptr, c_int, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: which, host_is_fp64
         type(c_ptr), value :: host
      end function
      integer(c_int) function wt_get_sections(ctx, ntf, dtf) bind(C, name="wt_get_sections")
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: ntf(*)
         real(c_double), intent(out) :: dtf(*)
      end function
      integer(c_int) function wt_get_tube_lengths(ctx, length) bind(C, name="wt_get_tube_lengths")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: length(*)
      end function
      integer(c_int) function wt_stage_times(ctx, ms, cap) bind(C, name="wt_stage_times")
         import :: c_ptr, c_int, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: ms(*)
         integer(c_int32_t), value :: cap
      end function
   end interface

contains

   ! Message of the last error of ctx (or of wt_create when ctx is null) as a Fortran string.
   function wt_last_error_str(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n
      interface
         integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
         end function
      end interface
      p = wt_last_error(ctx)
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function

   ! Stop with the library's message if rc signals an error.
   subroutine wt_check(ctx, rc, where)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= 0) then
         write (*, '(a,a,a,i0,a,a)') 'wovengpu error in ', where, ' (', rc, '): ', wt_last_error_str(ctx)
         error stop 2
      end if
   end subroutine
end module wovengpu
