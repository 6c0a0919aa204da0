!! fft_hip_tests.F90 -- the reference's own FFT and transpose checks, run in
!! Fortran through the decomp_2d / decomp_2d_fft modules over libd2dhip at
!! one rank (no MPI).  Mirrors:
!!   examples/fft_physical_x/fft_c2c_x.f90 and fft_r2c_x.f90 (and the _z
!!     copies): analytic inputs, ntest forward+backward pairs with 1/N,
!!     mean |err| per point <= 50 * epsilon * ntest;
!!   examples/fft_multiple_grids/fft_multiple_grids.f90: three grids plus a
!!     local engine object, decomp_2d_init on (nx+1, ny+1, nz+1);
!!   examples/test2d/test2d.f90: index field through the four transposes,
!!     exact compare;
!!   examples/dtt/fftw_f03_dtt_x.f90 and _z.f90: every DCT / DST type triple
!!     on 8 x 9 x 10, forward then backward, normalised, Linf <= 100 epsilon
!!     on the transformed points;
!!   examples/halo_test/halo_test.f90: a divergence through halo cells in
!!     each pencil against the one from transposes, within 2 epsilon, and
!!     the halo cells of a periodic grid against the wrapped pencil.
!! Exits with a non-zero status (decomp_2d_abort / error stop) on failure.
program fft_hip_tests

   use, intrinsic :: iso_c_binding
   use decomp_2d_constants
   use decomp_2d
   use decomp_2d_fft
   use d2dhip_c, only: hipMalloc, hipFree, hipMemcpy, HIP_H2D, HIP_D2H
   use m_halo
   use decomp_2d_profiler

   implicit none

   integer, parameter :: nx = 17, ny = 13, nz = 11, ntest = 10
   type(decomp_2d_fft_engine) :: local_engine
   integer :: igrid

   ! the library's regions on (transpose, fft, d2d; no io), plus one of our own
   call decomp_profiler_prep([.true., .false., .true., .true.])
   if (decomp_profiler_d2d) call decomp_profiler_start("fft_hip_tests")
   call decomp_2d_init(nx + 1, ny + 1, nz + 1, 0, 0)
   call check_transposes()

   ! single grid, both formats (fft_c2c_x / fft_r2c_x / fft_*_z)
   call decomp_2d_fft_init(PHYSICAL_IN_X, nx, ny, nz)
   call ntest_c2c(ntest, nx, ny, nz, 'c2c X')
   call ntest_r2c(ntest, nx, ny, nz, 'r2c X')
   call decomp_2d_fft_finalize
   call decomp_2d_fft_init(PHYSICAL_IN_Z, nx, ny, nz)
   call ntest_c2c(ntest, nx, ny, nz, 'c2c Z')
   call ntest_r2c(ntest, nx, ny, nz, 'r2c Z')
   call check_device_path(nx, ny, nz)
   call decomp_2d_fft_finalize

   ! multiple grids (fft_multiple_grids.f90:48-66)
   call decomp_2d_fft_set_ngrid(3)
   call decomp_2d_fft_init(PHYSICAL_IN_X, nx, ny, nz, 1)
   call decomp_2d_fft_init(PHYSICAL_IN_Z, nx, ny, nz, 2)
   call decomp_2d_fft_init(PHYSICAL_IN_X, nx, ny + 2, nz + 16, 3)
   call local_engine%init(PHYSICAL_IN_Z, nx + 16, ny + 2, nz)
   do igrid = 1, 3
      call decomp_2d_fft_use_grid(igrid)
      call grid_tests(ntest, igrid)
   end do
   call check_batch(3)
   call local_engine%use_it()
   call ntest_c2c(ntest, nx + 16, ny + 2, nz, 'c2c local engine')
   call ntest_r2c(ntest, nx + 16, ny + 2, nz, 'r2c local engine')
   call local_engine%fin()
   call decomp_2d_fft_finalize
   call check_dtt(PHYSICAL_IN_X)
   call check_dtt(PHYSICAL_IN_Z)
   call check_halo()
   call decomp_2d_finalize
   call decomp_2d_init(nx, ny, nz, 0, 0, periodic_bc=[.true., .true., .true.])
   call check_halo_periodic()
   call decomp_2d_finalize
   if (decomp_profiler_d2d) call decomp_profiler_end("fft_hip_tests")
   call decomp_profiler_fin()
   write (*, *) 'fft_hip_tests completed'

contains

   subroutine grid_tests(nt, g)
      integer, intent(in) :: nt, g
      type(decomp_info), pointer :: ph
      character(len=32) :: tag
      ph => decomp_2d_fft_get_ph()
      write (tag, '(a,i0)') 'grid ', g
      call ntest_c2c(nt, ph%xsz(1), ph%xsz(2), ph%xsz(3), trim(tag)//' c2c')
      call ntest_r2c(nt, ph%xsz(1), ph%xsz(2), ph%xsz(3), trim(tag)//' r2c')
   end subroutine grid_tests

   ! examples/fft_physical_x/fft_c2c_x.f90:64-154
   subroutine ntest_c2c(nt, mx, my, mz, tag)
      integer, intent(in) :: nt, mx, my, mz
      character(len=*), intent(in) :: tag
      complex(mytype), allocatable, dimension(:, :, :) :: in, out
      real(mytype) :: dr, di, error
      integer :: i, j, k, m
      allocate (in(mx, my, mz), out(mx, my, mz))
      do k = 1, mz
         do j = 1, my
            do i = 1, mx
               dr = real(i, mytype) / real(mx, mytype) * real(j, mytype) &
                    / real(my, mytype) * real(k, mytype) / real(mz, mytype)
               in(i, j, k) = cmplx(dr, dr, mytype)
            end do
         end do
      end do
      do m = 1, nt
         call decomp_2d_fft_3d(in, out, DECOMP_2D_FFT_FORWARD)
         call decomp_2d_fft_3d(out, in, DECOMP_2D_FFT_BACKWARD)
         in = in / (real(mx, mytype) * real(my, mytype) * real(mz, mytype))
      end do
      error = 0._mytype
      do k = 1, mz
         do j = 1, my
            do i = 1, mx
               dr = real(i, mytype) / real(mx, mytype) * real(j, mytype) &
                    / real(my, mytype) * real(k, mytype) / real(mz, mytype)
               di = dr - aimag(in(i, j, k))
               dr = dr - real(in(i, j, k), mytype)
               error = error + sqrt(dr * dr + di * di)
            end do
         end do
      end do
      error = error / (real(mx, mytype) * real(my, mytype) * real(mz, mytype))
      write (*, *) tag, ' error / mesh point: ', error
      if (error > epsilon(error) * 50 * nt) &
         call decomp_2d_abort(__FILE__, __LINE__, int(log10(error)), tag)
      deallocate (in, out)
   end subroutine ntest_c2c

   ! examples/fft_physical_x/fft_r2c_x.f90:66-152
   subroutine ntest_r2c(nt, mx, my, mz, tag)
      integer, intent(in) :: nt, mx, my, mz
      character(len=*), intent(in) :: tag
      real(mytype), allocatable, dimension(:, :, :) :: in_r
      complex(mytype), allocatable, dimension(:, :, :) :: out
      integer, dimension(3) :: fft_start, fft_end, fft_size
      real(mytype) :: dr, error
      integer :: i, j, k, m
      call decomp_2d_fft_get_size(fft_start, fft_end, fft_size)
      allocate (in_r(mx, my, mz))
      allocate (out(fft_size(1), fft_size(2), fft_size(3)))
      do k = 1, mz
         do j = 1, my
            do i = 1, mx
               in_r(i, j, k) = real(i, mytype) / real(mx, mytype) * real(j, mytype) &
                               / real(my, mytype) * real(k, mytype) / real(mz, mytype)
            end do
         end do
      end do
      do m = 1, nt
         call decomp_2d_fft_3d(in_r, out)
         call decomp_2d_fft_3d(out, in_r)
         in_r = in_r / (real(mx, mytype) * real(my, mytype) * real(mz, mytype))
      end do
      error = 0._mytype
      do k = 1, mz
         do j = 1, my
            do i = 1, mx
               dr = real(i, mytype) / real(mx, mytype) * real(j, mytype) &
                    / real(my, mytype) * real(k, mytype) / real(mz, mytype)
               error = error + abs(in_r(i, j, k) - dr)
            end do
         end do
      end do
      error = error / (real(mx, mytype) * real(my, mytype) * real(mz, mytype))
      write (*, *) tag, ' error / mesh point: ', error
      if (error > epsilon(error) * 50 * nt) &
         call decomp_2d_abort(__FILE__, __LINE__, int(log10(error)), tag)
      deallocate (in_r, out)
   end subroutine ntest_r2c

   ! examples/test2d/test2d.f90:77-199: u = i + (j-1)nx + (k-1)nx*ny through
   ! x -> y -> z -> y -> x, exact.
   subroutine check_transposes()
      real(mytype), allocatable, dimension(:, :, :) :: u1, u2, u3
      integer :: i, j, k
      call alloc_x(u1, decomp_main, .true.)
      call alloc_y(u2, decomp_main, .true.)
      call alloc_z(u3, decomp_main, .true.)
      do k = xstart(3), xend(3)
         do j = xstart(2), xend(2)
            do i = xstart(1), xend(1)
               u1(i, j, k) = real(i + (j - 1) * nx_global + (k - 1) * nx_global * ny_global, mytype)
            end do
         end do
      end do
      call transpose_x_to_y(u1, u2)
      call transpose_y_to_z(u2, u3)
      call transpose_z_to_y(u3, u2)
      call transpose_y_to_x(u2, u1)
      do k = xstart(3), xend(3)
         do j = xstart(2), xend(2)
            do i = xstart(1), xend(1)
               if (u1(i, j, k) /= real(i + (j - 1) * nx_global + (k - 1) * nx_global * ny_global, mytype)) &
                  call decomp_2d_abort(__FILE__, __LINE__, 10, "transpose test")
            end do
         end do
      end do
      write (*, *) 'transposes exact'
   end subroutine check_transposes

   ! Device-resident data (the GPU backend's use_device path): r2c then c2r
   ! on hipMalloc'ed buffers, round trip to 1/N.
   subroutine check_device_path(mx, my, mz)
      integer, intent(in) :: mx, my, mz
      real(mytype), allocatable, target :: a(:, :, :), b(:, :, :)
      type(c_ptr) :: d_r, d_c
      integer(c_size_t) :: nr, nc
      integer, dimension(3) :: fs, fe, fz
      real(mytype) :: err
      integer(c_int) :: st
      allocate (a(mx, my, mz), b(mx, my, mz))
      call random_number(a)
      call decomp_2d_fft_get_size(fs, fe, fz)
      nr = 8_c_size_t * mx * my * mz
      nc = 16_c_size_t * fz(1) * fz(2) * fz(3)
      st = hipMalloc(d_r, nr)
      st = ior(st, hipMalloc(d_c, nc))
      st = ior(st, hipMemcpy(d_r, c_loc(a), nr, HIP_H2D))
      if (st /= 0) call decomp_2d_abort(__FILE__, __LINE__, int(st), "hip")
      call decomp_2d_fft_3d_r2c_dev(d_r, d_c)
      call decomp_2d_fft_3d_c2r_dev(d_c, d_r)
      st = hipMemcpy(c_loc(b), d_r, nr, HIP_D2H)
      err = maxval(abs(b / (real(mx, mytype) * my * mz) - a))
      write (*, *) 'device path max error: ', err
      if (err > 1e-13_mytype) call decomp_2d_abort(__FILE__, __LINE__, 11, "device path")
      st = hipFree(d_r)
      st = hipFree(d_c)
   end subroutine check_device_path

   ! all grids' c2c transforms in one batched launch against per-grid calls
   subroutine check_batch(ng)
      integer, intent(in) :: ng
      type(c_ptr) :: d_in(ng), d_out(ng), d_ref(ng)
      integer(c_size_t) :: nb(ng)
      complex(mytype), allocatable, target :: a(:), b(:), c(:)
      real(mytype), allocatable :: re(:), im(:)
      type(decomp_info), pointer :: ph
      integer :: g, sgn
      integer(c_int) :: st
      real(mytype) :: err
      do g = 1, ng
         call decomp_2d_fft_use_grid(g)
         ph => decomp_2d_fft_get_ph()
         nb(g) = 16_c_size_t * ph%xsz(1) * ph%xsz(2) * ph%xsz(3)
         allocate (re(product(ph%xsz)), im(product(ph%xsz)))
         call random_number(re)
         call random_number(im)
         a = cmplx(re, im, mytype)
         deallocate (re, im)
         st = hipMalloc(d_in(g), nb(g))
         st = ior(st, hipMalloc(d_out(g), nb(g)))
         st = ior(st, hipMalloc(d_ref(g), nb(g)))
         st = ior(st, hipMemcpy(d_in(g), c_loc(a), nb(g), HIP_H2D))
         if (st /= 0) call decomp_2d_abort(__FILE__, __LINE__, int(st), "hip")
      end do
      do sgn = -1, 1, 2
         call decomp_2d_fft_3d_batch_dev(d_in, d_out, sgn)
         do g = 1, ng
            call decomp_2d_fft_use_grid(g)
            call decomp_2d_fft_3d_dev(d_in(g), d_ref(g), sgn)
            allocate (b(nb(g) / 16), c(nb(g) / 16))
            st = hipMemcpy(c_loc(b), d_out(g), nb(g), HIP_D2H)
            st = ior(st, hipMemcpy(c_loc(c), d_ref(g), nb(g), HIP_D2H))
            if (st /= 0) call decomp_2d_abort(__FILE__, __LINE__, int(st), "hip")
            err = maxval(abs(b - c)) / maxval(abs(c))
            if (err > 1e-13_mytype) call decomp_2d_abort(__FILE__, __LINE__, 12, "batch vs grid")
            deallocate (b, c)
         end do
      end do
      ! forward then backward in one launch: out = N * in
      call decomp_2d_fft_3d_batch_dev(d_in, d_out, 0)
      do g = 1, ng
         allocate (b(nb(g) / 16), c(nb(g) / 16))
         st = hipMemcpy(c_loc(b), d_out(g), nb(g), HIP_D2H)
         st = ior(st, hipMemcpy(c_loc(c), d_in(g), nb(g), HIP_D2H))
         err = maxval(abs(b / real(nb(g) / 16, mytype) - c))
         write (*, *) 'batch grid ', g, ' round trip max error: ', err
         if (err > 1e-13_mytype) call decomp_2d_abort(__FILE__, __LINE__, 13, "batch round trip")
         deallocate (b, c)
         st = hipFree(d_in(g))
         st = hipFree(d_out(g))
         st = hipFree(d_ref(g))
      end do
   end subroutine check_batch

   ! examples/dtt/fftw_f03_dtt_x.f90:69-216 (and the _z copy) at one rank
   subroutine check_dtt(fmt)
      integer, intent(in) :: fmt
      integer, parameter :: mx = 8, my = 9, mz = 10
      type(decomp_2d_fft_engine), target :: e
      real(mytype), allocatable :: a(:, :, :), b(:, :, :), ref(:, :, :)
      integer :: j, k, l, ii, jj, kk, ncase
      integer :: t(3), n(3), st(3), en(3)
      real(mytype) :: fac, errli
      allocate (a(mx, my, mz), b(mx, my, mz), ref(mx, my, mz))
      do kk = 1, mz
         do jj = 1, my
            do ii = 1, mx
               ref(ii, jj, kk) = real(ii, mytype) / real(mx, mytype) * real(jj, mytype) &
                                 / real(my, mytype) * real(kk, mytype) / real(mz, mytype)
            end do
         end do
      end do
      n = [mx, my, mz]
      ncase = 0
      do j = 1, 8
         do k = 1, 8
            do l = 1, 8
               t = [j, k, l]
               call e%init(fmt, mx, my, mz, opt_DTT=t)
               call e%use_it()
               if (any(decomp_2d_fft_get_dtt_input() /= t)) &
                  call decomp_2d_abort(__FILE__, __LINE__, 1, "DTT types not kept")
               a = ref
               b = 0._mytype
               call decomp_2d_dtt_3d_r2r(a, b, DECOMP_2D_FFT_FORWARD)
               call decomp_2d_dtt_3d_r2r(b, a, DECOMP_2D_FFT_BACKWARD)
               fac = 1._mytype
               do ii = 1, 3
                  if (t(ii) == 1 .or. t(ii) == 4) then
                     fac = fac * real(2 * n(ii) - 2, mytype)
                  else if (t(ii) == 5) then
                     fac = fac * real(2 * n(ii) + 2 - 2 * e%dtt(6 + ii), mytype)
                  else
                     fac = fac * real(2 * n(ii) - 2 * e%dtt(6 + ii), mytype)
                  end if
                  st(ii) = e%dtt(3 + ii)
                  en(ii) = n(ii) - e%dtt(6 + ii) + e%dtt(3 + ii) - 1
               end do
               a = a / fac
               errli = maxval(abs(a(st(1):en(1), st(2):en(2), st(3):en(3)) &
                                  - ref(st(1):en(1), st(2):en(2), st(3):en(3))))
               if (errli > 100 * epsilon(errli)) then
                  write (*, *) "check case ", t, e%dtt, errli
                  call decomp_2d_abort(__FILE__, __LINE__, 1, "DTT: Error above limit")
               end if
               call e%fin()
               ncase = ncase + 1
            end do
         end do
      end do
      ! the device-pointer form gives the host form's result bit for bit
      call e%init(fmt, mx, my, mz, opt_DTT=[2, 7, 4])
      call e%use_it()
      call check_dtt_dev(ref, mx, my, mz)
      call e%fin()
      write (*, '(a,i0,a,i0)') ' dtt checks passed: ', ncase, ' type triples, format ', fmt
      deallocate (a, b, ref)
   end subroutine check_dtt

   ! examples/halo_test/halo_test.f90:154-480 at one rank (non-periodic)
   subroutine check_halo()
      real(mytype), allocatable :: u(:, :, :), v(:, :, :), w(:, :, :), div(:, :, :), dh(:, :, :)
      real(mytype), allocatable :: uh(:, :, :), vh(:, :, :), wh(:, :, :)
      integer :: i, j, k, ip, s1, s2, s3
      real(mytype) :: err, mag
      s1 = xsize(1); s2 = xsize(2); s3 = xsize(3)
      allocate (u(s1, s2, s3), v(s1, s2, s3), w(s1, s2, s3), div(s1, s2, s3), dh(s1, s2, s3))
      do k = 1, s3
         do j = 1, s2
            do i = 1, s1
               u(i, j, k) = sin(0.3_mytype * i + 0.7_mytype * j) * cos(0.2_mytype * k)
               v(i, j, k) = cos(0.5_mytype * i) * sin(0.9_mytype * j + 0.1_mytype * k)
               w(i, j, k) = sin(0.4_mytype * i * j) + cos(0.6_mytype * k)
            end do
         end do
      end do
      ! the reference divergence (test_div_transpose: transposes are copies here)
      div = 0._mytype
      do k = 1, s3
         do j = 1, s2
            do i = 2, s1 - 1
               div(i, j, k) = u(i + 1, j, k) - u(i - 1, j, k)
            end do
         end do
      end do
      do k = 1, s3
         do j = 2, s2 - 1
            div(:, j, k) = div(:, j, k) + v(:, j + 1, k) - v(:, j - 1, k)
         end do
      end do
      do k = 2, s3 - 1
         div(:, :, k) = div(:, :, k) + w(:, :, k + 1) - w(:, :, k - 1)
      end do
      do ip = 1, 3
         dh = 0._mytype
         if (ip == 1) then  ! test_div_haloX
            call update_halo(v, vh, 1, opt_pencil=1)
            call update_halo(w, wh, 1, opt_pencil=1)
            if (any(shape(vh) /= [s1, s2 + 2, s3 + 2])) error stop "halo X size"
            do k = 1, s3
               do j = 1, s2
                  do i = 2, s1 - 1
                     dh(i, j, k) = (u(i + 1, j, k) - u(i - 1, j, k)) + (vh(i, j + 1, k) - vh(i, j - 1, k)) &
                                   + (wh(i, j, k + 1) - wh(i, j, k - 1))
                  end do
               end do
            end do
            deallocate (vh, wh)
         else if (ip == 2) then  ! test_div_haloY
            call update_halo(u, uh, 1, opt_pencil=2)
            call update_halo(w, wh, 1, opt_pencil=2)
            if (any(shape(uh) /= [s1 + 2, s2, s3 + 2])) error stop "halo Y size"
            do k = 1, s3
               do j = 2, s2 - 1
                  do i = 1, s1
                     dh(i, j, k) = (uh(i + 1, j, k) - uh(i - 1, j, k)) + (v(i, j + 1, k) - v(i, j - 1, k)) &
                                   + (wh(i, j, k + 1) - wh(i, j, k - 1))
                  end do
               end do
            end do
            deallocate (uh, wh)
         else  ! test_div_haloZ
            call update_halo(u, uh, 1, opt_pencil=3)
            call update_halo(v, vh, 1, opt_pencil=3)
            if (any(shape(uh) /= [s1 + 2, s2 + 2, s3])) error stop "halo Z size"
            do k = 2, s3 - 1
               do j = 1, s2
                  do i = 1, s1
                     dh(i, j, k) = (uh(i + 1, j, k) - uh(i - 1, j, k)) + (vh(i, j + 1, k) - vh(i, j - 1, k)) &
                                   + (w(i, j, k + 1) - w(i, j, k - 1))
                  end do
               end do
            end do
            deallocate (uh, vh)
         end if
         ! check_err: interior points only
         err = sqrt(sum((dh(2:s1 - 1, 2:s2 - 1, 2:s3 - 1) - div(2:s1 - 1, 2:s2 - 1, 2:s3 - 1))**2))
         mag = sqrt(sum(div(2:s1 - 1, 2:s2 - 1, 2:s3 - 1)**2))
         if (.not. (err < 2._mytype * epsilon(mag) * mag)) then
            write (*, *) "halo divergence, pencil ", ip, err, mag
            call decomp_2d_abort(__FILE__, __LINE__, ip, "Error in halo_test")
         end if
      end do
      write (*, *) 'halo divergence checks passed'
      deallocate (u, v, w, div, dh)
   end subroutine check_halo

   ! periodic_bc = all: the halo cells are the wrapped pencil (global indices too)
   subroutine check_halo_periodic()
      real(mytype), allocatable :: a(:, :, :), ah(:, :, :)
      integer :: s1, s2, s3, i, j, k
      type(halo_extents_t) :: h
      s1 = xsize(1); s2 = xsize(2); s3 = xsize(3)
      allocate (a(s1, s2, s3))
      do k = 1, s3
         do j = 1, s2
            do i = 1, s1
               a(i, j, k) = real(i + 100 * j + 10000 * k, mytype)
            end do
         end do
      end do
      call update_halo(a, ah, 2, opt_global=.true., opt_pencil=1)
      if (lbound(ah, 2) /= -1 .or. ubound(ah, 3) /= s3 + 2) error stop "halo global bounds"
      do k = -1, s3 + 2
         do j = -1, s2 + 2
            if (any(ah(:, j, k) /= a(:, modulo(j - 1, s2) + 1, modulo(k - 1, s3) + 1))) &
               call decomp_2d_abort(__FILE__, __LINE__, 1, "periodic halo cell")
         end do
      end do
      deallocate (ah)
      ! exchange_test's form: the halo array filled by hand, halo_exchange in place
      h = halo_extents_t(3, [s1, s2, s3], decomp_main, 1, .false.)
      allocate (ah(h%xs:h%xe, h%ys:h%ye, h%zs:h%ze))
      ah = -1._mytype
      ah(1:s1, 1:s2, :) = a
      call halo_exchange(ah, 3, h, 1, [s1, s2, s3])
      do j = 0, s2 + 1
         do i = 0, s1 + 1
            if (any(ah(i, j, :) /= a(modulo(i - 1, s1) + 1, modulo(j - 1, s2) + 1, :))) &
               call decomp_2d_abort(__FILE__, __LINE__, 3, "periodic halo cell")
         end do
      end do
      write (*, *) 'periodic halo checks passed'
      deallocate (a, ah)
   end subroutine check_halo_periodic

   subroutine check_dtt_dev(ref, mx, my, mz)
      integer, intent(in) :: mx, my, mz
      real(mytype), intent(in), target :: ref(mx, my, mz)
      real(mytype), allocatable, target :: h(:, :, :), hd(:, :, :)
      type(c_ptr) :: da, db
      integer(c_size_t) :: nb
      allocate (h(mx, my, mz), hd(mx, my, mz))
      h = ref
      call decomp_2d_dtt_3d_r2r(h, hd, DECOMP_2D_FFT_FORWARD)
      nb = 8_c_size_t * mx * my * mz
      da = c_null_ptr
      db = c_null_ptr
      if (hipMalloc(da, nb) /= 0 .or. hipMalloc(db, nb) /= 0) error stop "hipMalloc"
      if (hipMemcpy(da, c_loc(ref), nb, HIP_H2D) /= 0) error stop "hipMemcpy"
      call decomp_2d_dtt_3d_r2r_dev(da, db, DECOMP_2D_FFT_FORWARD)
      if (hipMemcpy(c_loc(h), db, nb, HIP_D2H) /= 0) error stop "hipMemcpy"
      if (any(h /= hd)) call decomp_2d_abort(__FILE__, __LINE__, 1, "DTT device form differs")
      if (hipFree(da) /= 0 .or. hipFree(db) /= 0) error stop "hipFree"
      deallocate (h, hd)
   end subroutine check_dtt_dev

end program fft_hip_tests
