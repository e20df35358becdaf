! iso_c_binding view of the engine's C ABI (include/ccgpu.h) for Fortran hosts.
! Scalars by value, arrays as assumed-size actuals, complex numbers as
! complex(c_double_complex) (layout-identical to the C side's interleaved doubles).
module ccgpu_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CCGPU_PNM = 0, CCGPU_SNM = 1
  integer(c_int), parameter :: CCGPU_INT_MINNESOTA = 0, CCGPU_INT_CHIRAL = 1
  integer(c_int), parameter :: CCGPU_PREC_FP64 = 0, CCGPU_PREC_FP32 = 1, CCGPU_PREC_V32 = 2, CCGPU_PREC_BF16X3 = 3
  integer(c_int), parameter :: CCGPU_BLOCK_HHHH = 1, CCGPU_BLOCK_HHHP = 2, CCGPU_BLOCK_HHPP = 3, &
       CCGPU_BLOCK_HPHP = 4, CCGPU_BLOCK_HPPP = 5, CCGPU_BLOCK_PPPP = 6

  type, bind(C) :: ccgpu_lattice_params
     integer(c_int) :: A
     integer(c_int) :: matter
     real(c_double) :: rho
     integer(c_int) :: nmax
     real(c_double) :: twist(3)
  end type ccgpu_lattice_params

  type, bind(C) :: ccgpu_dims
     integer(c_int) :: nsp, nholes, nparticles, nchannels, nph_channels, max_npp, max_nhh
     integer(c_long_long) :: t2_unique, t2_ordered, vpppp_unique, ring_elems
     real(c_double) :: box, volume, kf
     real(c_double) :: pp_ladder_flops, hh_ladder_flops, ring_flops
  end type ccgpu_dims

  type, bind(C) :: ccgpu_solve_opts
     integer(c_int) :: subspace, diis_step
     real(c_double) :: tol
     integer(c_int) :: maxit, precision, refine_every, use_graph, verbose
     real(c_double) :: fp32_switch
     integer(c_int) :: ccdt1 = 0, accel = 0
     real(c_double) :: mix = 0.d0
     integer(c_int) :: tail = 0, confirm = 0
  end type ccgpu_solve_opts

  type, bind(C) :: ccgpu_solve_result
     integer(c_int) :: iterations, converged
     real(c_double) :: e0(2), embpt2(2), eccd(2), ecorr_last(2), seconds
     integer(c_int) :: fp32_iterations, status
  end type ccgpu_solve_result

  interface
     integer(c_int) function ccgpu_plan(p, d) bind(C, name='ccgpu_plan')
       import :: c_int, ccgpu_lattice_params, ccgpu_dims
       type(ccgpu_lattice_params), intent(in) :: p
       type(ccgpu_dims), intent(out) :: d
     end function ccgpu_plan
     integer(c_int) function ccgpu_create(p, h) bind(C, name='ccgpu_create')
       import :: c_int, ccgpu_lattice_params
       type(ccgpu_lattice_params), intent(in) :: p
       integer(c_int), intent(out) :: h
     end function ccgpu_create
     integer(c_int) function ccgpu_destroy(h) bind(C, name='ccgpu_destroy')
       import :: c_int
       integer(c_int), value :: h
     end function ccgpu_destroy
     integer(c_int) function ccgpu_get_dims(h, d) bind(C, name='ccgpu_get_dims')
       import :: c_int, ccgpu_dims
       integer(c_int), value :: h
       type(ccgpu_dims), intent(out) :: d
     end function ccgpu_get_dims
     integer(c_int) function ccgpu_generate_interaction(h, kind) bind(C, name='ccgpu_generate_interaction')
       import :: c_int
       integer(c_int), value :: h, kind
     end function ccgpu_generate_interaction
     integer(c_int) function ccgpu_set_chiral(h, order, cutoff, ci) bind(C, name='ccgpu_set_chiral')
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: h, order
       real(c_double), value :: cutoff
       type(c_ptr), value :: ci
     end function ccgpu_set_chiral
     integer(c_int) function ccgpu_set_3nf(h, approx, ce, cd, cutoff, nexp) bind(C, name='ccgpu_set_3nf')
       import :: c_int, c_double
       integer(c_int), value :: h, approx, nexp
       real(c_double), value :: ce, cd, cutoff
     end function ccgpu_set_3nf
     integer(c_int) function ccgpu_get_e0(h, e0) bind(C, name='ccgpu_get_e0')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e0
     end function ccgpu_get_e0
     integer(c_int) function ccgpu_mbpt2_start(h, e) bind(C, name='ccgpu_mbpt2_start')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e
     end function ccgpu_mbpt2_start
     integer(c_int) function ccgpu_residual(h, prec) bind(C, name='ccgpu_residual')
       import :: c_int
       integer(c_int), value :: h, prec
     end function ccgpu_residual
     integer(c_int) function ccgpu_update(h, count, nstep, subspace, diis_step) bind(C, name='ccgpu_update')
       import :: c_int
       integer(c_int), value :: h, subspace, diis_step
       integer(c_int), intent(inout) :: count, nstep
     end function ccgpu_update
     integer(c_int) function ccgpu_energy(h, e) bind(C, name='ccgpu_energy')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e
     end function ccgpu_energy
     integer(c_int) function ccgpu_batch_solve(n, hs, o, r) bind(C, name='ccgpu_batch_solve')
       import :: c_int, ccgpu_solve_opts, ccgpu_solve_result
       integer(c_int), value :: n
       integer(c_int), intent(in) :: hs(*)
       type(ccgpu_solve_opts), intent(in) :: o
       type(ccgpu_solve_result), intent(out) :: r(*)
     end function ccgpu_batch_solve
     integer(c_int) function ccgpu_solve(h, o, r) bind(C, name='ccgpu_solve')
       import :: c_int, ccgpu_solve_opts, ccgpu_solve_result
       integer(c_int), value :: h
       type(ccgpu_solve_opts), intent(in) :: o
       type(ccgpu_solve_result), intent(out) :: r
     end function ccgpu_solve
     integer(c_int) function ccgpu_triples(h, e) bind(C, name='ccgpu_triples')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e
     end function ccgpu_triples
     integer(c_int) function ccgpu_triples_parts(h, e4, e5) bind(C, name='ccgpu_triples_parts')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e4, e5
     end function ccgpu_triples_parts
     integer(c_int) function ccgpu_setup_t3(h, e) bind(C, name='ccgpu_setup_t3')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e
     end function ccgpu_setup_t3
     integer(c_int) function ccgpu_t3_intermediate(h) bind(C, name='ccgpu_t3_intermediate')
       import :: c_int
       integer(c_int), value :: h
     end function ccgpu_t3_intermediate
     integer(c_int) function ccgpu_t3_energy(h, e) bind(C, name='ccgpu_t3_energy')
       import :: c_int, c_double_complex
       integer(c_int), value :: h
       complex(c_double_complex), intent(out) :: e
     end function ccgpu_t3_energy
     integer(c_int) function ccgpu_last_error(buf, n) bind(C, name='ccgpu_last_error')
       import :: c_int, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: n
     end function ccgpu_last_error
  end interface

contains

  ! Stop with the library's message when a call failed (the reference stops too,
  ! e.g. ccm_modules.f90:685-688, but here the message says why).
  subroutine ccgpu_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(1024)
    character(len=1024) :: msg
    integer :: n, k
    if (status == 0) return
    n = ccgpu_last_error(buf, 1024)
    msg = ' '
    do k = 1, min(max(n, 0), 1023)
       msg(k:k) = buf(k)
    end do
    write(0,'(a,a,a,a)') 'ccgpu error in ', what, ': ', trim(msg)
    error stop 2
  end subroutine ccgpu_check

end module ccgpu_iface
