! Drop-in Fortran host of the woven-turbulence generator on one MI355X.
!
! Reads ./centerline_input.dat and ./parameters_input.dat exactly as the reference program
! does (tube.f90:168-223, 341-379), runs the woven-field path on the GPU through the wovengpu
! C ABI, and writes the reference's outputs:
!   output/parameters_output.dat, output/centerline_output.dat   (tube.f90:171-258, 348-371)
!   output/test_data.dat       per tube ntf_average / ntf_max       (tube.f90:673-676)
!   output/stat/spectrum_vor_temp.dat  spectrum of raw vorticity    (tube.f90:972-978)
!   output/stat/spectrum.dat   spectrum of normalised velocity      (tube.f90:1158-1164)
!   output/stat/pdf_um.dat, pdf_wm.dat  PDFs of u and omega, 1000 bins (tube.f90:1166-1185)
!   output/stat/Sp.dat         structure functions, orders 1..10    (subroutine.f90:64-129)
!   output/field_wh.dat        |omega| and omega.u, downsampled, raw FP32 (tube.f90:1055-1073)
!   output/info.dat            summary                              (tube.f90:1197-1211)
! Usage: weave_host [32|64]   (compute precision, default 64 = the reference's arithmetic)
program weave_host
   use iso_c_binding
   use wovengpu
   implicit none

   integer, parameter :: Sp_order = 10, bin_num = 1000
   integer :: nline, npointall, i, j, k, i_line, nx, n_d, nek, nsep, r_d, nb, prec, ios
   integer(c_int32_t), allocatable :: npointlist(:), n_T_line(:), if_filter_line(:), seps(:)
   integer(c_int32_t), allocatable :: ntf(:)
   real(c_double), allocatable :: cxall(:), cyall(:), czall(:), sigma_line(:), Gamma_line(:)
   real(c_double), allocatable :: E(:), Sp(:), dtf(:)
   real(c_double), allocatable, target :: f(:, :, :, :)
   real(c_float), allocatable :: data_box(:, :, :, :)
   real(c_float) :: sigma_out, Gamma_out
   real(c_float) :: vari(bin_num), pdf(bin_num)
   type(wt_params) :: p
   type(wt_seg_stats) :: st
   type(c_ptr) :: ctx
   real(c_double) :: etot, up, enstrophy, lengthtot, dr
   real(c_double) :: ms(32)
   real(c_double), allocatable :: tube_len(:)
   integer(8) :: clk_rate, clk_total0, clk0, clk1
   real(8) :: t_init, t_total, t_post, t_spline, t_constr, share
   integer(c_int) :: rc
   integer(c_int64_t) :: sec0
   character(len=32) :: arg
   real(8), parameter :: pi2 = 8.0d0*datan(1.0d0)

   prec = 64
   if (command_argument_count() >= 1) then
      call get_command_argument(1, arg)
      read (arg, *, iostat=ios) prec
      if (ios /= 0 .or. (prec /= 32 .and. prec /= 64)) error stop 'precision must be 32 or 64'
   end if
   call execute_command_line('mkdir -p output/stat', wait=.true.)

   ! ---- centerline_input.dat (tube.f90:168-177, 341-379) ----
   open (90, file='./centerline_input.dat', status='old', action='read')
   open (31, file='./output/centerline_output.dat', status='replace')
   read (90, *) nline
   write (31, *) nline
   allocate (npointlist(nline), sigma_line(nline), Gamma_line(nline), n_T_line(nline), if_filter_line(nline))
   do i = 1, nline
      read (90, *) npointlist(i)
      write (31, *) npointlist(i)
   end do
   npointall = sum(npointlist)
   allocate (cxall(npointall), cyall(npointall), czall(npointall))
   do i = 1, npointall
      read (90, *) cxall(i), cyall(i), czall(i)
      write (31, *) cxall(i), cyall(i), czall(i)
   end do
   close (90)
   close (31)

   ! ---- parameters_input.dat (tube.f90:191-223) ----
   open (512, file='./parameters_input.dat', status='old', action='read')
   read (512, *)
   read (512, *) p%N
   read (512, *)
   read (512, *) p%n_d
   read (512, *)
   read (512, *) p%line_s
   read (512, *)
   read (512, *) p%line_e
   read (512, *)
   read (512, *) p%sigma_ratio
   read (512, *)
   read (512, *) p%eta_t
   read (512, *)
   read (512, *)
   read (512, *) p%ntf_times
   read (512, *)
   read (512, *) p%ntf_pre
   read (512, *)
   read (512, *) p%Rtube_ratio
   read (512, *)
   read (512, *) p%dzeta_sup
   read (512, *)
   read (512, *)
   do i_line = 1, nline
      read (512, *) i, sigma_line(i), Gamma_line(i), n_T_line(i), if_filter_line(i)
   end do
   close (512)
   nx = p%N
   n_d = p%n_d

   open (19, file='./output/parameters_output.dat', status='replace')
   write (19, *) 'N'
   write (19, *) nx
   write (19, *) 'N_downsample'
   write (19, *) n_d
   write (19, *) 'line_s'
   write (19, *) p%line_s
   write (19, *) 'line_e'
   write (19, *) p%line_e
   write (19, *) 'sigma_ratio'
   write (19, *) p%sigma_ratio
   write (19, *) 'eta_t'
   write (19, *) p%eta_t
   write (19, *) '---------------'
   write (19, *) 'ntf_times'
   write (19, *) p%ntf_times
   write (19, *) 'ntf_pre'
   write (19, *) p%ntf_pre
   write (19, *) 'Rtube_ratio'
   write (19, *) p%Rtube_ratio
   write (19, *) 'dzeta_sup'
   write (19, *) p%dzeta_sup
   write (19, *) '---------------'
   write (19, *) 'variables="i_line","sigma","Gamma","n_T","if_filter"'
   do i_line = 1, nline
      sigma_out = real(sigma_line(i_line))
      Gamma_out = real(Gamma_line(i_line))
      write (19, *) i_line, sigma_out, Gamma_out, n_T_line(i_line), if_filter_line(i_line)
   end do
   close (19)

   ! ---- GPU path ----
   rc = wt_create(ctx, int(nx, c_int32_t), int(prec, c_int32_t), 0_c_int32_t)
   call wt_check(c_null_ptr, rc, 'wt_create')
   call system_clock(count_rate=clk_rate)
   call system_clock(clk_total0)
   call wt_check(ctx, wt_set_params(ctx, p), 'wt_set_params')
   call wt_check(ctx, wt_set_tubes(ctx, int(nline, c_int32_t), npointlist, sigma_line, Gamma_line, &
                                   n_T_line, if_filter_line, cxall, cyall, czall), 'wt_set_tubes')
   call system_clock(clk1)
   t_init = real(clk1 - clk_total0, 8)/real(clk_rate, 8)
   call wt_check(ctx, wt_build_segments(ctx, st), 'wt_build_segments')

   ! test_data.dat: per tube, integer ntf_average over its n_sec+1 sections and ntf_max
   allocate (ntf(st%nsec_all), dtf(st%nsec_all))
   call wt_check(ctx, wt_get_sections(ctx, ntf, dtf), 'wt_get_sections')
   open (519, file='./output/test_data.dat', status='replace')
   sec0 = 0
   do i_line = 1, nline
      nb = npointlist(i_line) + 1
      write (519, *) 'ntf_average=', sum(ntf(sec0 + 1:sec0 + nb))/nb
      write (519, *) 'ntf_max=', maxval(ntf(sec0 + 1:sec0 + nb))
      sec0 = sec0 + nb
   end do
   close (519)
   write (*, *) 'dtf_min =', st%dtf_min

   call wt_check(ctx, wt_bin(ctx), 'wt_bin')
   call wt_check(ctx, wt_deposit(ctx, c_null_ptr), 'wt_deposit')
   call system_clock(clk1)
   t_total = real(clk1 - clk_total0, 8)/real(clk_rate, 8)
   rc = wt_stage_times(ctx, ms, 32_c_int32_t)
   t_spline = ms(1)*1.0d-3
   t_constr = (ms(2) + ms(3))*1.0d-3
   allocate (tube_len(nline))
   call wt_check(ctx, wt_get_tube_lengths(ctx, tube_len), 'wt_get_tube_lengths')
   lengthtot = sum(tube_len)                        ! every tube, as tube.f90:502

   call system_clock(clk0)
   call wt_check(ctx, wt_biot_savart(ctx, etot, up, enstrophy), 'wt_biot_savart')

   nek = wt_nek(int(nx, c_int32_t))
   allocate (E(nek))
   rc = wt_spectrum(ctx, 0_c_int32_t, E, int(nek, c_int32_t))
   if (rc < 0) call wt_check(ctx, rc, 'wt_spectrum')
   open (499, file='./output/stat/spectrum_vor_temp.dat', status='replace')
   do i = 1, nek
      write (499, *) i, real(E(i))
   end do
   close (499)
   rc = wt_spectrum(ctx, 1_c_int32_t, E, int(nek, c_int32_t))
   if (rc < 0) call wt_check(ctx, rc, 'wt_spectrum')
   open (49, file='./output/stat/spectrum.dat', status='replace')
   do i = 1, nek
      write (49, *) i, real(E(i))
   end do
   close (49)

   ! pdf_um.dat, pdf_wm.dat (export_pdf, subroutine.f90:42-62)
   open (3201, file='./output/stat/pdf_um.dat', status='replace')
   call wt_check(ctx, wt_pdf(ctx, 1_c_int32_t, int(bin_num, c_int32_t), vari, pdf, c_null_ptr), 'wt_pdf')
   do i = 1, bin_num
      write (3201, *) vari(i), pdf(i)
   end do
   close (3201)
   open (3204, file='./output/stat/pdf_wm.dat', status='replace')
   call wt_check(ctx, wt_pdf(ctx, 0_c_int32_t, int(bin_num, c_int32_t), vari, pdf, c_null_ptr), 'wt_pdf')
   do i = 1, bin_num
      write (3204, *) vari(i), pdf(i)
   end do
   close (3204)

   ! Sp.dat (subroutine.f90:114-123)
   nsep = wt_sp_separations(int(nx, c_int32_t), seps, 0_c_int32_t)
   allocate (seps(nsep), Sp(nsep*Sp_order))
   nsep = wt_sp_separations(int(nx, c_int32_t), seps, int(nsep, c_int32_t))
   call wt_check(ctx, wt_structure_functions(ctx, 0_c_int32_t, 1_c_int32_t, int(Sp_order, c_int32_t), &
                                             int(nsep, c_int32_t), seps, Sp), 'wt_structure_functions')
   dr = pi2/nx
   open (3301, file='./output/stat/Sp.dat', status='replace')
   do j = 1, nsep
      write (3301, "(ES16.5)", advance='NO') seps(j)*dr
      do i = 1, Sp_order
         write (3301, "(ES16.5)", advance='NO') Sp((i - 1)*nsep + j)
      end do
      write (3301, *)
   end do
   close (3301)

   ! field_wh.dat: (|omega|, omega.u) at i*r_d, raw FP32 (tube.f90:1055-1073, subroutine.f90:1389-1399)
   allocate (f(nx, nx, nx, 6))
   do i = 1, 6
      call wt_check(ctx, wt_get_field(ctx, int(i - 1, c_int32_t), c_loc(f(1, 1, 1, i)), 1_c_int32_t), 'wt_get_field')
   end do
   r_d = nx/n_d
   nb = nx/r_d
   allocate (data_box(nb, nb, nb, 2))
   do k = 1, nb
      do j = 1, nb
         do i = 1, nb
            data_box(i, j, k, 1) = real(sqrt(f(i*r_d, j*r_d, k*r_d, 1)**2 + f(i*r_d, j*r_d, k*r_d, 2)**2 + &
                                             f(i*r_d, j*r_d, k*r_d, 3)**2))
            data_box(i, j, k, 2) = real(f(i*r_d, j*r_d, k*r_d, 1)*f(i*r_d, j*r_d, k*r_d, 4) + &
                                        f(i*r_d, j*r_d, k*r_d, 2)*f(i*r_d, j*r_d, k*r_d, 5) + &
                                        f(i*r_d, j*r_d, k*r_d, 3)*f(i*r_d, j*r_d, k*r_d, 6))
         end do
      end do
   end do
   open (unit=99, file='./output/field_wh.dat', access='stream', form='unformatted', status='replace')
   write (99) data_box
   close (99)

   call system_clock(clk1)
   t_post = real(clk1 - clk0, 8)/real(clk_rate, 8)

   ! info.dat in the reference's layout (tube.f90:519-525, 532, 756-757, 941-952, 983, 1022,
   ! 1033-1034, 1193-1207). All tubes go through the GPU in single launches, so a tube's spline and
   ! construction times are the launches' times (segment stage; binning + deposition) shared out by
   ! its section count. "Initialization time" is the upload of the tube table; the preprocessing
   ! kernels (tube.f90:392-515) run inside the segment stage.
   open (192, file='./output/info.dat', status='replace')
   write (192, *) 'Initialization time is', t_init
   write (192, *) '=============================='
   write (192, *) '   [i] [Section count] [Spline time] [Construction time]'
   do i_line = p%line_s, p%line_e
      share = real(npointlist(i_line), 8)/real(sum(npointlist(p%line_s:p%line_e)), 8)
      write (192, '(I6)', advance='no') i_line
      write (192, '(I16)', advance='no') npointlist(i_line)
      write (192, '(ES18.3)', advance='no') t_spline*share
      write (192, '(ES18.3)', advance='no') t_constr*share
      write (192, *)
   end do
   write (192, *) '=============================='
   write (192, *) 'Total construction time =', t_total
   write (192, *) 'velocity calculation begin'
   write (192, *) 'velocity calculation done'
   write (192, *) 'Etot_temp =', etot
   write (192, *) 'up_temp =', up
   write (192, *) 'Post processing time =', t_post
   write (192, *) '=============================='
   write (192, *) 'N = ', nx
   write (192, *) 'Total length of curve = ', lengthtot
   write (192, *) 'Gamma_1 = ', Gamma_line(1)/up
   write (192, *) 'twist eta = ', p%eta_t
   write (192, *) 'Etot = ', etot/up/up
   write (192, *) 'Enstrophy = ', enstrophy
   close (192)
   write (*, *) 'up =', up, ' Etot =', etot, ' Enstrophy =', enstrophy
   call wt_destroy(ctx)
end program weave_host
