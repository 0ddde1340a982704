! hipnegf_cudautils.f90 -- Fortran iso_c_binding interfaces of libhipnegf's
! block operations (include/hipnegf.h, section 1) under the names libNEGF's
! Fortran calls: the interface block of src/cudautils.F90:236-431 and the
! decimation interface of src/contselfenergy.F90:80-103, one for one, bound to
! the hipnegf_* symbols instead of the cu_* ones of gpuroutines.cu.
!
! Drop-in use: lib_param.F90 takes cublasHandle / cusolverDnHandle from this
! module (same bind(C) layout as lib_param.F90:90-96: one c_ptr), and
! cudautils.F90 / contselfenergy.F90 replace their interface blocks with
! `use hipnegf_cudautils`.  Link with -lhipnegf.
!
! Differences a caller sees:
!  * cu_cudaGetDeviceProperties is not provided (bound but never called by
!    libNEGF, and its Fortran interface does not match the C function);
!  * cudaDeviceSynchronize binds hipnegf_deviceSynchronize;
!  * SGFACC of cu_{C,Z}decimation is passed by value, as the C function
!    takes it (the reference's interface passes it by reference);
!  * singular inverses and non-converged decimations return 1 / 2 instead of
!    ending the process.
module hipnegf_cudautils
  use iso_c_binding
  implicit none
  private

  type, bind(C), public :: cublasHandle
    type(c_ptr) :: handle
  end type cublasHandle

  type, bind(C), public :: cusolverDnHandle
    type(c_ptr) :: handle
  end type cusolverDnHandle

  public :: cu_cublasInit, cu_cublasFinalize, cu_cusolverInit, cu_cusolverFinalize
  public :: cudaDeviceSynchronize, cu_cudaFreeAsync, cu_cudaGetDeviceCount
  public :: cu_cudaMallocAsync, cu_cudaGetDevice, cu_cudaSetDevice
  public :: cu_createMat, cu_copyMatD2H, cu_copyMatH2D, cu_copyMatH2D_async
  public :: cu_deleteMat, cu_meminfo
  public :: cu_CmultMat, cu_ZmultMat, cu_Cinverse, cu_Zinverse, cu_Cmatsum, cu_Zmatsum
  public :: cu_Cinitmat, cu_Zinitmat, cu_Ctrace, cu_Ztrace, cu_Cmatcopy, cu_Zmatcopy
  public :: cu_Casum, cu_Zasum, cu_Cdecimation, cu_Zdecimation

  interface
    ! ---- handles: one hipnegf context stands in for both library handles
    integer(c_int) function cu_cublasInit(hcublas) bind(C, name='hipnegf_blasInit')
      import :: c_int, cublasHandle
      type(cublasHandle) :: hcublas
    end function cu_cublasInit

    integer(c_int) function cu_cublasFinalize(hcublas) bind(C, name='hipnegf_blasFinalize')
      import :: c_int, cublasHandle
      type(cublasHandle), value :: hcublas
    end function cu_cublasFinalize

    integer(c_int) function cu_cusolverInit(hcusolver) bind(C, name='hipnegf_solverInit')
      import :: c_int, cusolverDnHandle
      type(cusolverDnHandle) :: hcusolver
    end function cu_cusolverInit

    integer(c_int) function cu_cusolverFinalize(hcusolver) bind(C, name='hipnegf_solverFinalize')
      import :: c_int, cusolverDnHandle
      type(cusolverDnHandle), value :: hcusolver
    end function cu_cusolverFinalize

    integer(c_int) function cudaDeviceSynchronize() bind(C, name='hipnegf_deviceSynchronize')
      import :: c_int
    end function cudaDeviceSynchronize

    ! ---- device memory (pointers passed by reference are set or nulled)
    integer(c_int) function cu_cudaFreeAsync(d_A) bind(C, name='hipnegf_freeAsync')
      import :: c_int, c_ptr
      type(c_ptr) :: d_A
    end function cu_cudaFreeAsync

    integer(c_int) function cu_cudaGetDeviceCount(count) bind(C, name='hipnegf_getDeviceCount')
      import :: c_int
      integer(c_int) :: count
    end function cu_cudaGetDeviceCount

    integer(c_int) function cu_cudaMallocAsync(d_A, siz) bind(C, name='hipnegf_mallocAsync')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr) :: d_A
      integer(c_size_t), value :: siz
    end function cu_cudaMallocAsync

    integer(c_int) function cu_cudaGetDevice(device) bind(C, name='hipnegf_getDevice')
      import :: c_int
      integer(c_int) :: device
    end function cu_cudaGetDevice

    integer(c_int) function cu_cudaSetDevice(device) bind(C, name='hipnegf_setDevice')
      import :: c_int
      integer(c_int), value :: device
    end function cu_cudaSetDevice

    integer(c_int) function cu_createMat(d_A, siz) bind(C, name='hipnegf_createMat')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr) :: d_A
      integer(c_size_t), value :: siz
    end function cu_createMat

    integer(c_int) function cu_copyMatD2H(h_A, d_A, siz) bind(C, name='hipnegf_copyMatD2H')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h_A, d_A
      integer(c_size_t), value :: siz
    end function cu_copyMatD2H

    integer(c_int) function cu_copyMatH2D(h_A, d_A, siz) bind(C, name='hipnegf_copyMatH2D')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h_A, d_A
      integer(c_size_t), value :: siz
    end function cu_copyMatH2D

    integer(c_int) function cu_copyMatH2D_async(h_A, d_A, siz) &
        & bind(C, name='hipnegf_copyMatH2D_async')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h_A, d_A
      integer(c_size_t), value :: siz
    end function cu_copyMatH2D_async

    integer(c_int) function cu_deleteMat(d_A) bind(C, name='hipnegf_deleteMat')
      import :: c_int, c_ptr
      type(c_ptr) :: d_A
    end function cu_deleteMat

    integer(c_int) function cu_meminfo(freemem, totalmem) bind(C, name='hipnegf_meminfo')
      import :: c_int, c_int64_t
      integer(c_int64_t) :: freemem, totalmem
    end function cu_meminfo

    ! ---- complex float (C) and complex double (Z) block operations
    integer(c_int) function cu_CmultMat(hcublas, m, n, k, alpha, d_A, d_B, beta, d_C, dagger) &
        & bind(C, name='hipnegf_CmultMat')
      import :: c_int, c_size_t, c_ptr, c_float_complex, cublasHandle
      type(cublasHandle), value :: hcublas
      integer(c_size_t), value :: m, n, k
      complex(c_float_complex) :: alpha, beta
      type(c_ptr), value :: d_A, d_B, d_C
      integer(c_int), value :: dagger
    end function cu_CmultMat

    integer(c_int) function cu_ZmultMat(hcublas, m, n, k, alpha, d_A, d_B, beta, d_C, dagger) &
        & bind(C, name='hipnegf_ZmultMat')
      import :: c_int, c_size_t, c_ptr, c_double_complex, cublasHandle
      type(cublasHandle), value :: hcublas
      integer(c_size_t), value :: m, n, k
      complex(c_double_complex) :: alpha, beta
      type(c_ptr), value :: d_A, d_B, d_C
      integer(c_int), value :: dagger
    end function cu_ZmultMat

    integer(c_int) function cu_Cinverse(hcublas, hcusolver, d_A, d_Ainv, N) &
        & bind(C, name='hipnegf_Cinverse')
      import :: c_int, c_size_t, c_ptr, cublasHandle, cusolverDnHandle
      type(cublasHandle), value :: hcublas
      type(cusolverDnHandle), value :: hcusolver
      type(c_ptr), value :: d_A, d_Ainv
      integer(c_size_t), value :: N
    end function cu_Cinverse

    integer(c_int) function cu_Zinverse(hcublas, hcusolver, d_A, d_Ainv, N) &
        & bind(C, name='hipnegf_Zinverse')
      import :: c_int, c_size_t, c_ptr, cublasHandle, cusolverDnHandle
      type(cublasHandle), value :: hcublas
      type(cusolverDnHandle), value :: hcusolver
      type(c_ptr), value :: d_A, d_Ainv
      integer(c_size_t), value :: N
    end function cu_Zinverse

    integer(c_int) function cu_Cmatsum(hcublas, m, n, alpha, d_A, beta, d_B, d_C, dagger) &
        & bind(C, name='hipnegf_Cmatsum')
      import :: c_int, c_size_t, c_ptr, c_float_complex, cublasHandle
      type(cublasHandle), value :: hcublas
      integer(c_size_t), value :: m, n
      complex(c_float_complex) :: alpha, beta
      type(c_ptr), value :: d_A, d_B, d_C
      integer(c_int), value :: dagger
    end function cu_Cmatsum

    integer(c_int) function cu_Zmatsum(hcublas, m, n, alpha, d_A, beta, d_B, d_C, dagger) &
        & bind(C, name='hipnegf_Zmatsum')
      import :: c_int, c_size_t, c_ptr, c_double_complex, cublasHandle
      type(cublasHandle), value :: hcublas
      integer(c_size_t), value :: m, n
      complex(c_double_complex) :: alpha, beta
      type(c_ptr), value :: d_A, d_B, d_C
      integer(c_int), value :: dagger
    end function cu_Zmatsum

    integer(c_int) function cu_Cinitmat(d_A, nrow) bind(C, name='hipnegf_Cinitmat')
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: d_A
      integer(c_size_t), value :: nrow
    end function cu_Cinitmat

    integer(c_int) function cu_Zinitmat(d_A, nrow) bind(C, name='hipnegf_Zinitmat')
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: d_A
      integer(c_size_t), value :: nrow
    end function cu_Zinitmat

    ! h_tun: host array of nrow default logicals (4 bytes), read as int
    real(c_float) function cu_Ctrace(hcublas, d_A, nrow, h_tun, mask_present) &
        & bind(C, name='hipnegf_Ctrace')
      import :: c_int, c_size_t, c_ptr, c_float, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A, h_tun
      integer(c_size_t), value :: nrow
      integer(c_int), value :: mask_present
    end function cu_Ctrace

    real(c_double) function cu_Ztrace(hcublas, d_A, nrow, h_tun, mask_present) &
        & bind(C, name='hipnegf_Ztrace')
      import :: c_int, c_size_t, c_ptr, c_double, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A, h_tun
      integer(c_size_t), value :: nrow
      integer(c_int), value :: mask_present
    end function cu_Ztrace

    integer(c_int) function cu_Cmatcopy(hcublas, d_A, d_B, msize) bind(C, name='hipnegf_Cmatcopy')
      import :: c_int, c_size_t, c_ptr, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A, d_B
      integer(c_size_t), value :: msize
    end function cu_Cmatcopy

    integer(c_int) function cu_Zmatcopy(hcublas, d_A, d_B, msize) bind(C, name='hipnegf_Zmatcopy')
      import :: c_int, c_size_t, c_ptr, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A, d_B
      integer(c_size_t), value :: msize
    end function cu_Zmatcopy

    integer(c_int) function cu_Casum(hcublas, d_A, summ, N) bind(C, name='hipnegf_Casum')
      import :: c_int, c_size_t, c_ptr, c_float, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A
      real(c_float) :: summ
      integer(c_size_t), value :: N
    end function cu_Casum

    integer(c_int) function cu_Zasum(hcublas, d_A, summ, N) bind(C, name='hipnegf_Zasum')
      import :: c_int, c_size_t, c_ptr, c_double, cublasHandle
      type(cublasHandle), value :: hcublas
      type(c_ptr), value :: d_A
      real(c_double) :: summ
      integer(c_size_t), value :: N
    end function cu_Zasum

    ! host matrices; ncyc out; SGFACC by value (see the header of this file)
    integer(c_int) function cu_Cdecimation(hcublas, hcusolver, h_Go_out, h_Ao_in, h_Bo_in, &
        & h_Co_in, n, tf32, ncyc, SGFACC) bind(C, name='hipnegf_Cdecimation')
      import :: c_int, c_size_t, c_ptr, c_float, cublasHandle, cusolverDnHandle
      type(cublasHandle), value :: hcublas
      type(cusolverDnHandle), value :: hcusolver
      type(c_ptr), value :: h_Go_out, h_Ao_in, h_Bo_in, h_Co_in
      integer(c_size_t), value :: n
      integer(c_int), value :: tf32
      type(c_ptr), value :: ncyc
      real(c_float), value :: SGFACC
    end function cu_Cdecimation

    integer(c_int) function cu_Zdecimation(hcublas, hcusolver, h_Go_out, h_Ao_in, h_Bo_in, &
        & h_Co_in, n, tf32, ncyc, SGFACC) bind(C, name='hipnegf_Zdecimation')
      import :: c_int, c_size_t, c_ptr, c_double, cublasHandle, cusolverDnHandle
      type(cublasHandle), value :: hcublas
      type(cusolverDnHandle), value :: hcusolver
      type(c_ptr), value :: h_Go_out, h_Ao_in, h_Bo_in, h_Co_in
      integer(c_size_t), value :: n
      integer(c_int), value :: tf32
      type(c_ptr), value :: ncyc
      real(c_double), value :: SGFACC
    end function cu_Zdecimation
  end interface
end module hipnegf_cudautils
