! ccm_gpu: drop-in driver for the reference program ccm.exe on one GPU.
!
! Reads the reference's ccm_in deck (positional list-directed reads of
! ccm_main.f90:31-32 and 893-900), builds the lattice, interaction, Fock matrix
! and MBPT2 start through the C ABI, runs the ccd_iter loop (ccm_iter.f90:60-135)
! with one library call per reference routine (t2_intermediate, diis_t1t2,
! ccd_energy_save), optionally (T), and prints the reference's result block
! (ccm_main.f90:322-348).
!
! One optional line after the reference deck sets engine options:
!   ! interaction, DIIS subspace, DIIS step, tolerance, max iterations, precision, fp32 switch
!   chiral, 6, 6, 1e-6, 100, fp64, 1.0
! The interaction is chiral (the Delta-full chiral NN of the deck's first line,
! chiral order and cutoff, as the reference) or minnesota.
! With precision fp32 (or v32: FP32-stored pppp elements, FP64 arithmetic) the contractions run in FP32 up to the first DIIS
! extrapolation that follows an iteration with |dE| <= the switch (MeV; 0: never
! there) or until the FP32 |dE| meets the tolerance, and in FP64 afterwards
! (DIIS history kept), so the stopping test is always an FP64 one.
! Without it the reference constants are used (DIIS 6/6, 1e-6, 100 iterations).
program ccm_gpu
  use iso_c_binding
  use ccgpu_iface
  implicit none
  character(len=512) :: inputfile
  character(len=100) :: calc_type, input_type, boundary_conditions, cc_approx, interaction, precname
  integer :: order, below_ef, ntwist, nmax, my_twists, tnf_approx, nexp, ios
  integer :: subspace, diis_step, maxit, itimes, prec
  integer(c_int) :: h, count, nstep, st
  real(8) :: lambda, ce, cd, input, cutoff3, dens, kf, tol4r, fp32_switch, prev_dener
  logical :: fp32_phase, extrapolated, at_boundary
  real(4) :: tol4
  real(8), parameter :: pi = 3.14159265358979d0
  logical :: tnf_switch
  complex(c_double_complex) :: e0, ener1, ener2, dener, et, mbpt2, eccsd, eccsdt
  type(ccgpu_lattice_params) :: p
  type(ccgpu_dims) :: dims
  real(8) :: t_start, t_end

  call get_command_argument(1, inputfile)
  if (len_trim(inputfile) == 0) then
     write(0,*) 'usage: ccm_gpu ccm_in'
     error stop 1
  end if
  open(5, file=trim(inputfile), status='old', access='sequential')
  read(5,*); read(5,*) order, lambda
  read(5,*); read(5,*) ce, cd
  read(5,*); read(5,*) below_ef
  read(5,*); read(5,*) calc_type, input_type
  read(5,*); read(5,*) boundary_conditions
  read(5,*); read(5,*) input, ntwist, nmax, my_twists
  read(5,*); read(5,*) cc_approx
  read(5,*); read(5,*) tnf_switch, tnf_approx
  read(5,*); read(5,*) cutoff3, nexp
  interaction = 'chiral'
  subspace = 6; diis_step = 6; tol4 = 1.0e-6; maxit = 100; precname = 'fp64'; fp32_switch = 1.d0
  read(5,*,iostat=ios)
  if (ios == 0) read(5,*,iostat=ios) interaction, subspace, diis_step, tol4, maxit, precname, fp32_switch
  close(5)

  if (trim(interaction) /= 'minnesota' .and. trim(interaction) /= 'chiral') then
     write(0,'(2a)') 'ccm_gpu: unknown interaction ', trim(interaction)
     error stop 3
  end if
  if (tnf_switch .and. trim(interaction) /= 'chiral') then
     write(0,'(a)') 'ccm_gpu: the three-nucleon force belongs to the chiral interaction'
     error stop 3
  end if
  ! cE and cD vanish below NNLO (ccm_main.f90:33-36)
  if (order < 3) then
     ce = 0.d0
     cd = 0.d0
  end if
  if (trim(cc_approx) /= 'CCD' .and. trim(cc_approx) /= 'CCD(T)' .and. trim(cc_approx) /= 'mbpt2') then
     write(0,'(2a)') 'ccm_gpu: unsupported cluster approximation ', trim(cc_approx)
     error stop 3
  end if

  ! density / Fermi momentum (ccm_main.f90:924-951)
  select case (trim(calc_type))
  case ('pnm')
     p%matter = CCGPU_PNM
     if (trim(input_type) == 'density') then
        dens = input
     else
        kf = input
        dens = kf**3 / 3.d0 / pi**2
     end if
  case ('snm')
     p%matter = CCGPU_SNM
     if (trim(input_type) == 'density') then
        dens = input
     else
        kf = input
        dens = 2.d0 * kf**3 / 3.d0 / pi**2
     end if
  case default
     write(0,'(2a)') 'ccm_gpu: unknown matter ', trim(calc_type)
     error stop 3
  end select
  p%A = below_ef
  p%rho = dens
  p%nmax = nmax
  p%twist = 0.d0
  if (trim(boundary_conditions) == 'TABC') call twist_of(ntwist, my_twists, p%twist)

  call ccgpu_check(ccgpu_create(p, h), 'ccgpu_create')
  call ccgpu_check(ccgpu_get_dims(h, dims), 'ccgpu_get_dims')
  write(6,'(a,a)') 'Boundary Conditions ', trim(boundary_conditions)
  write(6,'(a,a)') 'CC approx ', trim(cc_approx)
  write(6,*) 'Boxsize', dims%box
  write(6,*) 'Density', dens
  write(6,*) 'K_f', dims%kf
  write(6,*) 'Nmax', nmax
  write(6,*) 'Total number of states', dims%nsp
  write(6,*) 'channels', dims%nchannels, 'ph channels', dims%nph_channels, 't2 unique', dims%t2_unique

  if (trim(interaction) == 'chiral') then
     ! delta_chiral_order and lambda_delta of the deck (ccm_main.f90:31, init_chp_constants)
     call ccgpu_check(ccgpu_set_chiral(h, int(order, c_int), real(lambda, c_double), c_null_ptr), 'ccgpu_set_chiral')
     ! tnf switch and approximation, 3NF regulator (ccm_main.f90:899-900): the
     ! normal ordering of normal_ordered_hamiltonian / setup_3nf_2bNO
     if (tnf_switch) call ccgpu_check(ccgpu_set_3nf(h, int(tnf_approx, c_int), real(ce, c_double), &
          real(cd, c_double), real(cutoff3, c_double), int(nexp, c_int)), 'ccgpu_set_3nf')
     call ccgpu_check(ccgpu_generate_interaction(h, CCGPU_INT_CHIRAL), 'ccgpu_generate_interaction')
  else
     call ccgpu_check(ccgpu_generate_interaction(h, CCGPU_INT_MINNESOTA), 'ccgpu_generate_interaction')
  end if
  call ccgpu_check(ccgpu_get_e0(h, e0), 'ccgpu_get_e0')
  write(6,*) 'Vacuum expectation value', e0

  ! setup_t_amplitudes + the first ccd_energy_save (ccm_iter.f90:18-19)
  call ccgpu_check(ccgpu_mbpt2_start(h, ener1), 'ccgpu_mbpt2_start')
  mbpt2 = e0 + ener1
  write(6,*) 'CCD energy', ener1 + e0
  eccsd = 0.d0
  eccsdt = 0.d0

  if (trim(cc_approx) /= 'mbpt2') then
     prec = CCGPU_PREC_FP64
     if (trim(precname) == 'fp32') prec = CCGPU_PREC_FP32
     if (trim(precname) == 'v32') prec = CCGPU_PREC_V32
     tol4r = real(tol4, 8)
     call cpu_time(t_start)
     count = 0
     nstep = 0
     ener2 = 0.d0
     dener = 4.0
     fp32_phase = prec /= CCGPU_PREC_FP64
     prev_dener = huge(1.d0)
     do itimes = 1, maxit
        if (abs(dener) > tol4r) then
           st = CCGPU_PREC_FP64
           if (fp32_phase) st = prec
           call ccgpu_check(ccgpu_residual(h, st), 'ccgpu_residual')
           count = count + 1
           nstep = nstep + 1
           call ccgpu_check(ccgpu_update(h, count, nstep, subspace, diis_step), 'ccgpu_update')
           extrapolated = count == 0
           call ccgpu_check(ccgpu_energy(h, ener2), 'ccgpu_energy')
           write(6,*) 'CCD energy', ener2 + e0
           dener = ener2 - ener1
           ener1 = ener2
           write(6,'(a,2i5,2e16.8)') 'dener', 0, itimes, real(dener)
           if (fp32_phase) then
              at_boundary = fp32_switch > 0.d0 .and. extrapolated .and. prev_dener <= fp32_switch
              if (at_boundary .or. abs(dener) <= tol4r) then
                 fp32_phase = .false.
                 ! an FP32 |dE| is no FP64 convergence test
                 if (abs(dener) <= 2.d0 * tol4r) dener = 2.d0 * tol4r
              end if
           end if
           if (.not. extrapolated) prev_dener = abs(dener)
        end if
     end do
     call cpu_time(t_end)
     eccsd = e0 + ener2
     write(6,*) 'Total execution time CCD', t_end - t_start
     if (trim(cc_approx) == 'CCD(T)') then
        call ccgpu_check(ccgpu_triples(h, et), 'ccgpu_triples')
        eccsdt = eccsd + et
        write(6,'(a,1x,2(F16.5,2x),a,1x,2(F16.5,3x))') 'Kf, Dens', dims%kf, dens, 'CCSD(T)', &
             real(et / below_ef), 0.d0
     end if
  end if

  ! result block (ccm_main.f90:322-348)
  write(6,*)
  write(6,*) '---------------------------------------------------------------------'
  write(6,*) 'Boxsize', dims%box, 'Nmax', nmax
  write(6,*) 'A, Density, Kf', below_ef, dens, dims%kf
  write(6,*) 'Twist averaged E0/N', e0 / below_ef
  write(6,*) 'Twist averaged mbpt2', mbpt2 / below_ef
  write(6,*) 'Twist averaged CCSD', eccsd / below_ef
  write(6,*)
  write(6,*) 'E/N', real(dims%kf), real(dens), real(e0 / below_ef), real(mbpt2 / below_ef), &
       real(eccsd / below_ef), real(eccsdt / below_ef)
  write(6,'(8f12.5)') ce, cd, real(dims%kf), real(dens), real(e0 / below_ef), &
       real(mbpt2 / below_ef), real(eccsd / below_ef), real(eccsdt / below_ef)
  write(6,*) '---------------------------------------------------------------------'
  call ccgpu_check(ccgpu_destroy(h), 'ccgpu_destroy')

contains

  ! Twist of run my_twists in a TABC deck: Gauss-Legendre nodes on [0, pi] per
  ! axis and the (nx <= ny <= nz) enumeration of ccm_main.f90:163-204, 909-921.
  subroutine twist_of(n, which, tw)
    integer, intent(in) :: n, which
    real(8), intent(out) :: tw(3)
    real(8) :: x(n), w(n)
    integer :: a, b, c, k
    call gauleg(0.d0, pi, x, w, n)
    k = 0
    tw = 0.d0
    do a = 1, n
       do b = a, n
          do c = b, n
             k = k + 1
             if (k == which) then
                tw(1) = x(a); tw(2) = x(b); tw(3) = x(c)
             end if
          end do
       end do
    end do
  end subroutine twist_of

  ! Gauss-Legendre nodes and weights on [x1, x2] by Newton iteration on P_n.
  subroutine gauleg(x1, x2, x, w, n)
    integer, intent(in) :: n
    real(8), intent(in) :: x1, x2
    real(8), intent(out) :: x(n), w(n)
    integer :: i, j, m
    real(8) :: p1, p2, p3, pp, xl, xm, z, z1
    m = (n + 1) / 2
    xm = 0.5d0 * (x2 + x1)
    xl = 0.5d0 * (x2 - x1)
    do i = 1, m
       z = cos(pi * (i - 0.25d0) / (n + 0.5d0))
       do
          p1 = 1.d0
          p2 = 0.d0
          do j = 1, n
             p3 = p2
             p2 = p1
             p1 = ((2.d0 * j - 1.d0) * z * p2 - (j - 1.d0) * p3) / j
          end do
          pp = n * (z * p1 - p2) / (z * z - 1.d0)
          z1 = z
          z = z1 - p1 / pp
          if (abs(z - z1) <= 1.d-14) exit
       end do
       x(i) = xm - xl * z
       x(n + 1 - i) = xm + xl * z
       w(i) = 2.d0 * xl / ((1.d0 - z * z) * pp * pp)
       w(n + 1 - i) = w(i)
    end do
  end subroutine gauleg

end program ccm_gpu
