! Calls every block operation of libhipnegf through the iso_c_binding
! interfaces of include/hipnegf_cudautils.f90, as libNEGF's cudautils.F90
! would, and checks each result on the host.  Prints one "ok <name>" line per
! check and "ALL OK" at the end; exits with status 1 on the first failure.
! Without a GPU it prints "NODEVICE" and exits 0 (link check only).
program blockops_check
  use iso_c_binding
  use hipnegf_cudautils
  implicit none
  type(cublasHandle) :: hb
  type(cusolverDnHandle) :: hs
  integer(c_int) :: rc, ndev, dev, ncyc
  integer(c_int64_t) :: fmem, tmem
  integer, parameter :: m = 37, n = 29, k = 45
  complex(c_double_complex), target :: A(m, k), Ad(k, m), B(k, n), Bd(n, k), C0(m, n), C(m, n)
  complex(c_double_complex), target :: S1(m, n), S2(m, n), St(n, m)
  complex(c_double_complex) :: alpha, beta, ref(m, n)
  type(c_ptr) :: dA, dB, dC, dX
  integer :: i, j

  rc = cu_cudaGetDeviceCount(ndev)
  if (rc /= 0 .or. ndev < 1) then
    print '(a)', 'NODEVICE'
    stop
  end if
  call check(cu_cudaGetDevice(dev) == 0 .and. dev >= 0, 'getDevice')
  call check(cu_cudaSetDevice(dev) == 0, 'setDevice')
  call check(cu_cublasInit(hb) == 0 .and. c_associated(hb%handle), 'cublasInit')
  call check(cu_cusolverInit(hs) == 0 .and. c_associated(hs%handle), 'cusolverInit')
  call check(cu_meminfo(fmem, tmem) == 0 .and. fmem > 0 .and. tmem >= fmem, 'meminfo')

  do j = 1, k
    do i = 1, m
      A(i, j) = cmplx(sin(0.37d0 * i + 0.11d0 * j), cos(0.23d0 * i * j), c_double_complex)
    end do
  end do
  do j = 1, n
    do i = 1, k
      B(i, j) = cmplx(cos(0.19d0 * i - 0.31d0 * j), sin(0.07d0 * i + 0.5d0 * j), c_double_complex)
    end do
  end do
  do j = 1, n
    do i = 1, m
      C0(i, j) = cmplx(0.01d0 * i, -0.02d0 * j, c_double_complex)
    end do
  end do
  Ad = conjg(transpose(A))
  Bd = conjg(transpose(B))
  alpha = (0.7d0, -0.3d0)
  beta = (-1.1d0, 0.4d0)

  ! ---- memory round trips
  call check(cu_createMat(dA, int(16 * m * k, c_size_t)) == 0, 'createMat')
  call check(cu_createMat(dB, int(16 * k * n, c_size_t)) == 0, 'createMat B')
  call check(cu_cudaMallocAsync(dC, int(16 * m * n, c_size_t)) == 0, 'cudaMallocAsync')
  call check(cudaDeviceSynchronize() == 0, 'deviceSynchronize')
  call check(cu_createMat(dX, int(16 * m * k, c_size_t)) == 0, 'createMat X')

  ! ---- ZmultMat, dagger 0 / 1 / 2
  call zup(A, dA, m * k)
  call zup(B, dB, k * n)
  call zup(C0, dC, m * n)
  call check(cu_ZmultMat(hb, int(m, c_size_t), int(n, c_size_t), int(k, c_size_t), alpha, dA, dB, &
      & beta, dC, 0) == 0, 'ZmultMat call')
  call zdown(C, dC, m * n)
  ref = alpha * matmul(A, B) + beta * C0
  call check(maxval(abs(C - ref)) <= 1d-12 * maxval(abs(ref)), 'ZmultMat N,N')
  call zup(Ad, dA, m * k)
  call zup(C0, dC, m * n)
  call check(cu_ZmultMat(hb, int(m, c_size_t), int(n, c_size_t), int(k, c_size_t), alpha, dA, dB, &
      & beta, dC, 1) == 0, 'ZmultMat call 1')
  call zdown(C, dC, m * n)
  call check(maxval(abs(C - ref)) <= 1d-12 * maxval(abs(ref)), 'ZmultMat A^H')
  call zup(A, dA, m * k)
  call zup(Bd, dB, k * n)
  call zup(C0, dC, m * n)
  call check(cu_ZmultMat(hb, int(m, c_size_t), int(n, c_size_t), int(k, c_size_t), alpha, dA, dB, &
      & beta, dC, 2) == 0, 'ZmultMat call 2')
  call zdown(C, dC, m * n)
  call check(maxval(abs(C - ref)) <= 1d-12 * maxval(abs(ref)), 'ZmultMat B^H')

  ! ---- Zmatsum, dagger 0 / 1 / 2 (m x n operands; A^H stored n x m)
  S1 = C0
  S2 = ref
  St = conjg(transpose(S1))
  call zup(S1, dA, m * n)
  call zup(S2, dB, m * n)
  call check(cu_Zmatsum(hb, int(m, c_size_t), int(n, c_size_t), alpha, dA, beta, dB, dC, 0) == 0, &
      & 'Zmatsum call')
  call zdown(C, dC, m * n)
  call check(maxval(abs(C - (alpha * S1 + beta * S2))) <= 1d-14 * maxval(abs(S2)), 'Zmatsum N')
  call zup(St, dA, m * n)
  call check(cu_Zmatsum(hb, int(m, c_size_t), int(n, c_size_t), alpha, dA, beta, dB, dC, 1) == 0, &
      & 'Zmatsum call 1')
  call zdown(C, dC, m * n)
  call check(maxval(abs(C - (alpha * S1 + beta * S2))) <= 1d-14 * maxval(abs(S2)), 'Zmatsum A^H')
  call zup(S1, dA, m * n)
  St = conjg(transpose(S2))
  call zup(St, dB, m * n)
  call check(cu_Zmatsum(hb, int(m, c_size_t), int(n, c_size_t), alpha, dA, beta, dB, dC, 2) == 0, &
      & 'Zmatsum call 2')
  call zdown(C, dC, m * n)
  call check(maxval(abs(C - (alpha * S1 + beta * S2))) <= 1d-14 * maxval(abs(S2)), 'Zmatsum B^H')

  ! ---- Zmatcopy, Zasum, Zinitmat, Ztrace (with and without the logical mask)
  call zup(A, dA, m * k)
  call check(cu_Zmatcopy(hb, dA, dX, int(m * k, c_size_t)) == 0, 'Zmatcopy call')
  call zdown_mk(dX)
  call zasum_check(dA)
  call ztrace_check()
  call zinverse_check(41)
  call zinverse_check(150)
  call zdecimation_check()
  call cops_check()

  call check(cu_deleteMat(dA) == 0 .and. .not. c_associated(dA), 'deleteMat')
  call check(cu_deleteMat(dB) == 0, 'deleteMat B')
  call check(cu_deleteMat(dX) == 0, 'deleteMat X')
  call check(cu_cudaFreeAsync(dC) == 0 .and. .not. c_associated(dC), 'cudaFreeAsync')
  call check(cudaDeviceSynchronize() == 0, 'deviceSynchronize 2')
  call check(cu_cusolverFinalize(hs) == 0, 'cusolverFinalize')
  call check(cu_cublasFinalize(hb) == 0, 'cublasFinalize')
  print '(a)', 'ALL OK'

contains

  subroutine check(ok, what)
    logical, intent(in) :: ok
    character(*), intent(in) :: what
    if (.not. ok) then
      print '(2a)', 'FAILED ', what
      error stop 1
    end if
    print '(2a)', 'ok ', what
  end subroutine check

  subroutine zup(h, d, nel)
    integer, intent(in) :: nel
    complex(c_double_complex), target, intent(in) :: h(*)
    type(c_ptr), intent(in) :: d
    call check(cu_copyMatH2D(c_loc(h), d, int(16 * nel, c_size_t)) == 0, 'copyMatH2D')
  end subroutine zup

  subroutine zdown(h, d, nel)
    integer, intent(in) :: nel
    complex(c_double_complex), target, intent(inout) :: h(*)
    type(c_ptr), intent(in) :: d
    call check(cu_copyMatD2H(c_loc(h), d, int(16 * nel, c_size_t)) == 0, 'copyMatD2H')
  end subroutine zdown

  subroutine zdown_mk(d)
    type(c_ptr), intent(in) :: d
    complex(c_double_complex), target :: X(m, k)
    call zdown(X, d, m * k)
    call check(all(X == A), 'Zmatcopy')
  end subroutine zdown_mk

  subroutine zasum_check(d)
    type(c_ptr), intent(in) :: d
    real(c_double) :: s
    call check(cu_Zasum(hb, d, s, int(m * k, c_size_t)) == 0, 'Zasum call')
    call check(abs(s - sum(abs(real(A)) + abs(aimag(A)))) <= 1d-12 * s, 'Zasum')
  end subroutine zasum_check

  subroutine ztrace_check()
    integer, parameter :: nt = 23
    complex(c_double_complex), target :: T(nt, nt)
    logical(c_int), target :: mask(nt)
    real(c_double) :: tr, ref1, ref2
    type(c_ptr) :: d
    integer :: ii
    do j = 1, nt
      do ii = 1, nt
        T(ii, j) = cmplx(0.1d0 * ii - 0.05d0 * j, 0.3d0 * ii, c_double_complex)
      end do
    end do
    ref1 = 0d0
    ref2 = 0d0
    do ii = 1, nt
      mask(ii) = mod(ii, 3) /= 0
      ref1 = ref1 + real(T(ii, ii))
      if (mask(ii)) ref2 = ref2 + real(T(ii, ii))
    end do
    call check(cu_createMat(d, int(16 * nt * nt, c_size_t)) == 0, 'createMat T')
    call zup(T, d, nt * nt)
    tr = cu_Ztrace(hb, d, int(nt, c_size_t), c_null_ptr, 0)
    call check(abs(tr - ref1) <= 1d-13 * abs(ref1), 'Ztrace')
    tr = cu_Ztrace(hb, d, int(nt, c_size_t), c_loc(mask), 1)
    call check(abs(tr - ref2) <= 1d-13 * abs(ref2), 'Ztrace masked')
    call check(cu_Zinitmat(d, int(nt, c_size_t)) == 0, 'Zinitmat call')
    call zdown(T, d, nt * nt)
    do j = 1, nt
      do ii = 1, nt
        if (ii == j) then
          if (T(ii, j) /= (1d0, 0d0)) call check(.false., 'Zinitmat diagonal')
        else
          if (T(ii, j) /= (0d0, 0d0)) call check(.false., 'Zinitmat off-diagonal')
        end if
      end do
    end do
    call check(.true., 'Zinitmat')
    call check(cu_deleteMat(d) == 0, 'deleteMat T')
  end subroutine ztrace_check

  subroutine zinverse_check(nn)
    integer, intent(in) :: nn
    complex(c_double_complex), target :: M(nn, nn), Mi(nn, nn), M2(nn, nn)
    complex(c_double_complex) :: P(nn, nn)
    type(c_ptr) :: d1, d2
    integer :: ii
    do j = 1, nn
      do ii = 1, nn
        M(ii, j) = cmplx(sin(1.3d0 * ii + 0.7d0 * j), cos(0.9d0 * ii - 0.4d0 * j), c_double_complex)
      end do
      M(j, j) = M(j, j) + cmplx(0.2d0 * sqrt(dble(nn)), 0.1d0, c_double_complex)
    end do
    call check(cu_createMat(d1, int(16 * nn * nn, c_size_t)) == 0, 'createMat M')
    call check(cu_createMat(d2, int(16 * nn * nn, c_size_t)) == 0, 'createMat Mi')
    call zup(M, d1, nn * nn)
    call check(cu_Zinverse(hb, hs, d1, d2, int(nn, c_size_t)) == 0, 'Zinverse call')
    call zdown(Mi, d2, nn * nn)
    call zdown(M2, d1, nn * nn)
    call check(all(M2 == M), 'Zinverse leaves A')
    P = matmul(M, Mi)
    do ii = 1, nn
      P(ii, ii) = P(ii, ii) - 1d0
    end do
    call check(maxval(abs(P)) <= 1d-11, 'Zinverse A*Ainv = I')
    call check(cu_deleteMat(d1) == 0 .and. cu_deleteMat(d2) == 0, 'deleteMat M')
  end subroutine zinverse_check

  ! diagonal chains: H00 = diag(eps_i), H01 = t I -> g_i of the 1-D chain
  subroutine zdecimation_check()
    integer, parameter :: nd = 3
    complex(c_double_complex), target :: Ao(nd, nd), Bo(nd, nd), Co(nd, nd), Go(nd, nd)
    complex(c_double_complex) :: z, w, r, g1, g2
    real(c_double) :: eps(nd), t
    integer(c_int), target :: nc
    integer :: ii
    eps = [0.3d0, -0.8d0, 1.1d0]
    t = -1d0
    z = cmplx(0.25d0, 1d-5, c_double_complex)
    Ao = 0; Bo = 0; Co = 0
    do ii = 1, nd
      Ao(ii, ii) = z - eps(ii)
      Bo(ii, ii) = -t
      Co(ii, ii) = -t
    end do
    call check(cu_Zdecimation(hb, hs, c_loc(Go), c_loc(Ao), c_loc(Bo), c_loc(Co), &
        & int(nd, c_size_t), 0, c_loc(nc), 1d-9 * nd * nd) == 0, 'Zdecimation call')
    call check(nc > 1 .and. nc < 300, 'Zdecimation ncyc')
    do ii = 1, nd
      w = z - eps(ii)
      r = sqrt(w * w - 4d0 * t * t)
      g1 = (w - r) / (2d0 * t * t)
      g2 = (w + r) / (2d0 * t * t)
      if (abs(t * g1) >= 1d0) g1 = g2
      if (abs(Go(ii, ii) - g1) > 1d-7 * abs(g1)) call check(.false., 'Zdecimation closed form')
    end do
    call check(abs(Go(1, 2)) + abs(Go(2, 3)) < 1d-12, 'Zdecimation decoupled')
    call check(.true., 'Zdecimation closed form')
  end subroutine zdecimation_check

  ! complex-float variants on the same data
  subroutine cops_check()
    integer, parameter :: nn = 19
    complex(c_float_complex), target :: X(nn, nn), Y(nn, nn), Z1(nn, nn), Xi(nn, nn)
    complex(c_float_complex) :: al, be
    complex(c_double_complex) :: Pd(nn, nn)
    complex(c_float_complex), target :: Ao(1, 1), Bo(1, 1), Co(1, 1), Go(1, 1)
    type(c_ptr) :: d1, d2, d3
    real(c_float) :: s, tr
    integer(c_int), target :: nc
    integer :: ii
    do j = 1, nn
      do ii = 1, nn
        X(ii, j) = cmplx(sin(0.5 * ii + 0.2 * j), cos(0.3 * ii * j), c_float_complex)
        Y(ii, j) = cmplx(cos(0.1 * ii - 0.6 * j), 0.25, c_float_complex)
      end do
      X(j, j) = X(j, j) + 3.0
    end do
    al = (1.5, 0.5)
    be = (-0.5, 2.0)
    call check(cu_createMat(d1, int(8 * nn * nn, c_size_t)) == 0, 'createMat c1')
    call check(cu_createMat(d2, int(8 * nn * nn, c_size_t)) == 0, 'createMat c2')
    call check(cu_createMat(d3, int(8 * nn * nn, c_size_t)) == 0, 'createMat c3')
    call check(cu_copyMatH2D(c_loc(X), d1, int(8 * nn * nn, c_size_t)) == 0, 'copy c1')
    call check(cu_copyMatH2D(c_loc(Y), d2, int(8 * nn * nn, c_size_t)) == 0, 'copy c2')
    call check(cu_Cinitmat(d3, int(nn, c_size_t)) == 0, 'Cinitmat')
    call check(cu_CmultMat(hb, int(nn, c_size_t), int(nn, c_size_t), int(nn, c_size_t), al, d1, d2, &
        & be, d3, 0) == 0, 'CmultMat call')
    call check(cu_copyMatD2H(c_loc(Z1), d3, int(8 * nn * nn, c_size_t)) == 0, 'copy back c3')
    Pd = al * matmul(cmplx(X, kind=c_double_complex), cmplx(Y, kind=c_double_complex))
    do ii = 1, nn
      Pd(ii, ii) = Pd(ii, ii) + be
    end do
    call check(maxval(abs(Z1 - Pd)) <= 1e-5 * maxval(abs(Pd)), 'CmultMat')
    call check(cu_Cmatsum(hb, int(nn, c_size_t), int(nn, c_size_t), al, d1, be, d2, d3, 0) == 0, &
        & 'Cmatsum call')
    call check(cu_copyMatD2H(c_loc(Z1), d3, int(8 * nn * nn, c_size_t)) == 0, 'copy back c3b')
    call check(maxval(abs(Z1 - (al * X + be * Y))) <= 1e-6 * maxval(abs(Z1)), 'Cmatsum')
    call check(cu_Cmatcopy(hb, d1, d3, int(nn * nn, c_size_t)) == 0, 'Cmatcopy call')
    call check(cu_copyMatD2H(c_loc(Z1), d3, int(8 * nn * nn, c_size_t)) == 0, 'copy back c3c')
    call check(all(Z1 == X), 'Cmatcopy')
    call check(cu_Casum(hb, d1, s, int(nn * nn, c_size_t)) == 0, 'Casum call')
    call check(abs(s - sum(abs(real(X)) + abs(aimag(X)))) <= 1e-5 * s, 'Casum')
    tr = cu_Ctrace(hb, d1, int(nn, c_size_t), c_null_ptr, 0)
    call check(abs(tr - sum([(real(X(ii, ii)), ii = 1, nn)])) <= 1e-5 * abs(tr), 'Ctrace')
    call check(cu_Cinverse(hb, hs, d1, d3, int(nn, c_size_t)) == 0, 'Cinverse call')
    call check(cu_copyMatD2H(c_loc(Xi), d3, int(8 * nn * nn, c_size_t)) == 0, 'copy back inv')
    Pd = matmul(cmplx(X, kind=c_double_complex), cmplx(Xi, kind=c_double_complex))
    do ii = 1, nn
      Pd(ii, ii) = Pd(ii, ii) - 1d0
    end do
    call check(maxval(abs(Pd)) <= 1e-4, 'Cinverse')
    Ao(1, 1) = cmplx(0.25 - 0.3, 1e-3, c_float_complex)
    Bo(1, 1) = (1.0, 0.0)
    Co(1, 1) = (1.0, 0.0)
    call check(cu_Cdecimation(hb, hs, c_loc(Go), c_loc(Ao), c_loc(Bo), c_loc(Co), 1_c_size_t, 0, &
        & c_loc(nc), 1e-6) == 0, 'Cdecimation call')
    Pd(1, 1) = cmplx(-0.05d0, 1d-3, c_double_complex)
    Pd(1, 2) = sqrt(Pd(1, 1) ** 2 - 4d0)
    Pd(2, 1) = (Pd(1, 1) - Pd(1, 2)) / 2d0
    if (abs(Pd(2, 1)) >= 1d0) Pd(2, 1) = (Pd(1, 1) + Pd(1, 2)) / 2d0
    call check(abs(Go(1, 1) - Pd(2, 1)) <= 1e-4 * abs(Pd(2, 1)), 'Cdecimation closed form')
    call check(cu_deleteMat(d1) == 0 .and. cu_deleteMat(d2) == 0 .and. cu_deleteMat(d3) == 0, &
        & 'deleteMat c')
  end subroutine cops_check

end program blockops_check
