! ccm_gpu: drop-in driver for the reference program ccm.exe on one GPU.
!
! Reads the reference's ccm_in deck (positional list-directed reads of
! ccm_main.f90:31-32 and 893-900), builds the lattice, interaction, Fock matrix
! and MBPT2 start through the C ABI, runs the ccd_iter loop (ccm_iter.f90:60-135)
! with one library call per reference routine (t2_intermediate, diis_t1t2,
! ccd_energy_save), optionally (T), and prints the reference's result block
! (ccm_main.f90:322-348).  cc_approx CCDT1 runs the reference's CCDT-1 branch
! of ccd_iter (ccm_iter.f90:26-92): the CCD iteration of T2 with the energy of
! the triples amplitude T3 = (W3 + W(T2))/D added every iteration.
!
! One optional line after the reference deck sets engine options:
!   ! interaction, DIIS subspace, DIIS step, tolerance, max iterations, precision, fp32 switch
!   chiral, 6, 6, 1e-6, 100, fp64, 1.0
! The interaction is chiral (the Delta-full chiral NN of the deck's first line,
! chiral order and cutoff, as the reference) or minnesota.
! With precision fp32 (or v32: FP32-stored pppp elements, FP64 arithmetic; or bf16x3: split
! bfloat16 products with FP32 sums) the contractions run in low precision up to the first DIIS
! extrapolation that follows an iteration with |dE| <= the switch (MeV; 0: never
! there) or until the FP32 |dE| meets the tolerance, and in FP64 afterwards
! (DIIS history kept), so the stopping test is always an FP64 one.
! Without it the reference constants are used (DIIS 6/6, 1e-6, 100 iterations).
!
! TABC decks: my_twists = k solves the k-th (nx <= ny <= nz) combination of
! the ntwist Gauss-Legendre twists on [0, pi] and prints its weighted E0 as the
! reference does (ccm_main.f90:163-245, where the MBPT2 and CCD sums are
! commented out and print as zero).  my_twists = 0 solves every combination in
! one batched solve and prints the weighted twist averages of E0, MBPT2, CCD
! (or CCDT-1) and CCD(T) (the accumulation the reference leaves commented out).
program ccm_gpu
  use iso_c_binding
  use ccgpu_iface
  implicit none
  character(len=512) :: inputfile
  character(len=100) :: calc_type, input_type, boundary_conditions, cc_approx, interaction, precname
  integer :: order, below_ef, ntwist, nmax, my_twists, tnf_approx, nexp, ios
  integer :: subspace, diis_step, maxit, itimes, prec
  real(8) :: tw_weight
  integer(c_int) :: h, count, nstep, st
  real(8) :: lambda, ce, cd, input, cutoff3, dens, kf, tol4r, fp32_switch, prev_dener
  logical :: fp32_phase, extrapolated, at_boundary
  real(4) :: tol4
  real(8), parameter :: pi = 3.14159265358979d0
  logical :: tnf_switch
  complex(c_double_complex) :: e0, ener1, ener2, dener, et, mbpt2, eccsd, eccsdt, e3, et4, et5
  logical :: ccdt1
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
  if (trim(cc_approx) /= 'CCD' .and. trim(cc_approx) /= 'CCD(T)' .and. trim(cc_approx) /= 'mbpt2' &
       .and. trim(cc_approx) /= 'CCDT1') then
     write(0,'(2a)') 'ccm_gpu: unsupported cluster approximation ', trim(cc_approx)
     error stop 3
  end if
  ccdt1 = trim(cc_approx) == 'CCDT1'
  if (ccdt1 .and. trim(interaction) /= 'chiral') then
     write(0,'(a)') 'ccm_gpu: CCDT1 needs the chiral three-nucleon force'
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
  prec = CCGPU_PREC_FP64
  if (trim(precname) == 'fp32') prec = CCGPU_PREC_FP32
  if (trim(precname) == 'v32') prec = CCGPU_PREC_V32
  if (trim(precname) == 'bf16x3') prec = CCGPU_PREC_BF16X3
  tw_weight = 1.d0
  if (trim(boundary_conditions) == 'TABC') then
     if (my_twists == 0) then
        call tabc_all()
        stop
     end if
     call twist_of(ntwist, my_twists, p%twist, tw_weight)
  end if

  call make_lattice(h)
  call ccgpu_check(ccgpu_get_dims(h, dims), 'ccgpu_get_dims')
  write(6,'(a,a)') 'Boundary Conditions ', trim(boundary_conditions)
  write(6,'(a,a)') 'CC approx ', trim(cc_approx)
  write(6,*) 'Boxsize', dims%box
  write(6,*) 'Density', dens
  write(6,*) 'K_f', dims%kf
  write(6,*) 'Nmax', nmax
  write(6,*) 'Total number of states', dims%nsp
  write(6,*) 'channels', dims%nchannels, 'ph channels', dims%nph_channels, 't2 unique', dims%t2_unique

  call ccgpu_check(ccgpu_get_e0(h, e0), 'ccgpu_get_e0')
  write(6,*) 'Vacuum expectation value', e0

  ! setup_t_amplitudes + the first ccd_energy_save (ccm_iter.f90:18-19)
  call ccgpu_check(ccgpu_mbpt2_start(h, ener1), 'ccgpu_mbpt2_start')
  mbpt2 = e0 + ener1
  write(6,*) 'CCD energy', ener1 + e0
  if (ccdt1) then
     ! setup_t3 and the first t3_energy (ccm_iter.f90:26-40)
     call ccgpu_check(ccgpu_setup_t3(h, e3), 'ccgpu_setup_t3')
     write(6,*) 'CCDT-1 energy', e3
     ener1 = ener1 + e3
  end if
  eccsd = 0.d0
  eccsdt = 0.d0

  if (trim(cc_approx) /= 'mbpt2') then
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
           if (ccdt1) call ccgpu_check(ccgpu_t3_intermediate(h), 'ccgpu_t3_intermediate')
           call ccgpu_check(ccgpu_residual(h, st), 'ccgpu_residual')
           count = count + 1
           nstep = nstep + 1
           call ccgpu_check(ccgpu_update(h, count, nstep, subspace, diis_step), 'ccgpu_update')
           extrapolated = count == 0
           call ccgpu_check(ccgpu_energy(h, ener2), 'ccgpu_energy')
           write(6,*) 'CCD energy', ener2 + e0
           if (ccdt1) then
              call ccgpu_check(ccgpu_t3_energy(h, e3), 'ccgpu_t3_energy')
              write(6,*) 'CCDT-1 energy', e3
              ener2 = ener2 + e3
           end if
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
        ! et_4, or et_5 at tnf_approx 3 (ccm_triples.f90:725-732)
        call ccgpu_check(ccgpu_triples(h, et), 'ccgpu_triples')
        call ccgpu_check(ccgpu_triples_parts(h, et4, et5), 'ccgpu_triples_parts')
        eccsdt = eccsd + et
        write(6,'(a,1x,2(F16.5,2x),a,1x,2(F16.5,3x))') 'Kf, Dens', dims%kf, dens, 'CCSD(T)', &
             real(et4 / below_ef), real(et5 / below_ef)
     end if
  end if

  ! TABC, one twist: the weighted E0 only (ccm_main.f90:238-245)
  if (trim(boundary_conditions) == 'TABC') then
     e0 = tw_weight * e0
     mbpt2 = 0.d0
     eccsd = 0.d0
     eccsdt = 0.d0
  end if
  call result_block(dims%box, dims%kf, e0, mbpt2, eccsd, eccsdt)
  call ccgpu_check(ccgpu_destroy(h), 'ccgpu_destroy')

contains

  ! The reference's result block (ccm_main.f90:318-348); hbarc and the
  ! single-precision nucleon mass of its constants module (ccm_modules.f90:91-93).
  subroutine result_block(box, kfv, e0v, mpv, ccv, ctv)
    real(c_double), intent(in) :: box, kfv
    complex(c_double_complex), intent(in) :: e0v, mpv, ccv, ctv
    real(8), parameter :: hbarc = 197.326968d0
    real(8), parameter :: p_mass = real(939.565, 8)
    write(6,*)
    write(6,*)
    write(6,*) '---------------------------------------------------------------------'
    write(6,*) 'Boxsize', box, 'Nmax', nmax
    write(6,*) 'A, Density, Kf', below_ef, dens, kfv
    write(6,*) 'Twist averaged E0/N', e0v / below_ef
    write(6,*) 'Twist averaged mbpt2', mpv / below_ef
    write(6,*) 'Twist averaged CCSD', ccv / below_ef
    write(6,*) 'kinetic_TF', 0.3d0 * (hbarc**2 * kfv**2 / p_mass) * below_ef
    write(6,*) dens, e0v / below_ef, ccv
    write(6,*)
    write(6,*) 'E/N', real(kfv), real(dens), real(e0v / below_ef), real(mpv / below_ef), &
         real(ccv / below_ef), real(ctv / below_ef)
    write(6,'(8f12.5)') ce, cd, real(kfv), real(dens), real(e0v / below_ef), &
         real(mpv / below_ef), real(ccv / below_ef), real(ctv / below_ef)
    write(6,*) '---------------------------------------------------------------------'
  end subroutine result_block

  ! Lattice p on the GPU with the deck's interaction.
  subroutine make_lattice(hh)
    integer(c_int), intent(out) :: hh
    call ccgpu_check(ccgpu_create(p, hh), 'ccgpu_create')
    if (trim(interaction) == 'chiral') then
       ! delta_chiral_order and lambda_delta of the deck (ccm_main.f90:31, init_chp_constants)
       call ccgpu_check(ccgpu_set_chiral(hh, int(order, c_int), real(lambda, c_double), c_null_ptr), &
            'ccgpu_set_chiral')
       ! tnf switch and approximation, 3NF regulator (ccm_main.f90:899-900): the
       ! normal ordering of normal_ordered_hamiltonian / setup_3nf_2bNO
       if (tnf_switch) then
          call ccgpu_check(ccgpu_set_3nf(hh, int(tnf_approx, c_int), real(ce, c_double), &
               real(cd, c_double), real(cutoff3, c_double), int(nexp, c_int)), 'ccgpu_set_3nf')
       else if (ccdt1) then
          ! not normal-ordered; the constants of the CCDT-1 three-body element
          call ccgpu_check(ccgpu_set_3nf(hh, -1_c_int, real(ce, c_double), &
               real(cd, c_double), real(cutoff3, c_double), int(nexp, c_int)), 'ccgpu_set_3nf')
       end if
       call ccgpu_check(ccgpu_generate_interaction(hh, CCGPU_INT_CHIRAL), 'ccgpu_generate_interaction')
    else
       call ccgpu_check(ccgpu_generate_interaction(hh, CCGPU_INT_MINNESOTA), 'ccgpu_generate_interaction')
    end if
  end subroutine make_lattice

  ! Symmetry factor of a twist combination (ccm_main.f90:236-240).
  real(8) function combo_factor(a, b, c)
    integer, intent(in) :: a, b, c
    if (a == b .and. a == c) then
       combo_factor = 1.d0
    else if (a == b .or. a == c .or. b == c) then
       combo_factor = 3.d0
    else
       combo_factor = 6.d0
    end if
  end function combo_factor

  ! Twist of run my_twists in a TABC deck and its weight: Gauss-Legendre nodes
  ! on [0, pi] per axis and the (nx <= ny <= nz) enumeration of
  ! ccm_main.f90:163-204, 909-921; weight factor w_x w_y w_z / pi^3 (242).
  subroutine twist_of(n, which, tw, wt)
    integer, intent(in) :: n, which
    real(8), intent(out) :: tw(3), wt
    real(8) :: x(n), w(n)
    integer :: a, b, c, k, ijk(3)
    call gauleg(0.d0, pi, x, w, n)
    k = 0
    ijk = 0
    tw = 0.d0
    wt = 0.d0
    do a = 1, n
       do b = a, n
          do c = b, n
             k = k + 1
             if (k == which) then
                tw(1) = x(a); tw(2) = x(b); tw(3) = x(c)
                wt = combo_factor(a, b, c) * w(a) * w(b) * w(c) / pi**3
                ijk = (/ a, b, c /)
             end if
          end do
       end do
    end do
    write(6,*) '# of Twist angle combinations', k
    if (k < which .or. which < 1) then
       write(0,'(a,i0,a,i0)') 'ccm_gpu: my_twists ', which, ' out of 1..', k
       error stop 3
    end if
    write(6,*) 'my twists are: ', ijk
    write(6,'(a,3es25.17)') ' twist angles', tw
  end subroutine twist_of

  ! Every twist combination of a TABC deck as one batched solve; weighted
  ! averages of E0, MBPT2, CCD and CCD(T).
  subroutine tabc_all()
    integer :: n, a, b, c, k, ncomb
    real(8) :: x(ntwist), w(ntwist)
    integer(c_int), allocatable :: hs(:)
    real(8), allocatable :: wts(:)
    complex(c_double_complex), allocatable :: e0k(:), mpk(:), cck(:), ctk(:)
    type(ccgpu_solve_result), allocatable :: res(:)
    type(ccgpu_solve_opts) :: o
    type(ccgpu_dims) :: d
    complex(c_double_complex) :: e1, etk, e0_av, mbpt2_av, eccsd_av, eccsdt_av
    n = ntwist
    call gauleg(0.d0, pi, x, w, n)
    ncomb = n * (n + 1) * (n + 2) / 6
    allocate(hs(ncomb), wts(ncomb), e0k(ncomb), mpk(ncomb), cck(ncomb), ctk(ncomb), res(ncomb))
    res(:)%iterations = 0
    k = 0
    do a = 1, n
       do b = a, n
          do c = b, n
             k = k + 1
             wts(k) = combo_factor(a, b, c) * w(a) * w(b) * w(c) / pi**3
             p%twist = (/ x(a), x(b), x(c) /)
             call make_lattice(hs(k))
             call ccgpu_check(ccgpu_get_e0(hs(k), e0k(k)), 'ccgpu_get_e0')
             call ccgpu_check(ccgpu_mbpt2_start(hs(k), e1), 'ccgpu_mbpt2_start')
             mpk(k) = e0k(k) + e1
             cck(k) = mpk(k)
             ctk(k) = 0.d0
          end do
       end do
    end do
    call ccgpu_check(ccgpu_get_dims(hs(1), d), 'ccgpu_get_dims')
    write(6,*) '# of Twist angle combinations', ncomb
    if (trim(cc_approx) /= 'mbpt2') then
       o%subspace = subspace
       o%diis_step = diis_step
       o%tol = real(tol4, 8)
       o%maxit = maxit
       o%precision = prec
       o%refine_every = 0
       o%use_graph = 1
       o%verbose = 0
       o%fp32_switch = fp32_switch
       o%ccdt1 = 0
       if (ccdt1) o%ccdt1 = 1
       call ccgpu_check(ccgpu_batch_solve(int(ncomb, c_int), hs, o, res), 'ccgpu_batch_solve')
       do k = 1, ncomb
          cck(k) = cmplx(res(k)%eccd(1), res(k)%eccd(2), c_double_complex)
          ctk(k) = cck(k)
          if (trim(cc_approx) == 'CCD(T)') then
             call ccgpu_check(ccgpu_triples(hs(k), etk), 'ccgpu_triples')
             ctk(k) = cck(k) + etk
          end if
       end do
    end if
    e0_av = 0.d0
    mbpt2_av = 0.d0
    eccsd_av = 0.d0
    eccsdt_av = 0.d0
    k = 0
    do a = 1, n
       do b = a, n
          do c = b, n
             k = k + 1
             write(6,'(a,i4,3es25.17,es25.17,3f16.8,i5)') 'twist', k, x(a), x(b), x(c), wts(k), &
                  real(e0k(k)) / below_ef, real(mpk(k)) / below_ef, real(cck(k)) / below_ef, res(k)%iterations
             e0_av = e0_av + wts(k) * e0k(k)
             mbpt2_av = mbpt2_av + wts(k) * mpk(k)
             eccsd_av = eccsd_av + wts(k) * cck(k)
             eccsdt_av = eccsdt_av + wts(k) * ctk(k)
             call ccgpu_check(ccgpu_destroy(hs(k)), 'ccgpu_destroy')
          end do
       end do
    end do
    if (trim(cc_approx) /= 'CCD(T)') eccsdt_av = 0.d0
    if (trim(cc_approx) == 'mbpt2') eccsd_av = 0.d0
    call result_block(d%box, d%kf, e0_av, mbpt2_av, eccsd_av, eccsdt_av)
  end subroutine tabc_all

  ! Gauss-Legendre nodes and weights on [x1, x2] with the arithmetic of the
  ! reference's gauss_legendre (ccm_modules.f90:924-959): initial guess from the
  ! literal 3.141592654, Newton steps until two iterates differ by <= 3e-14.
  ! The twists of a TABC deck, and through them the order of degenerate
  ! momenta, are then the reference's to the last bit.
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
       z1 = 0.d0
       z = cos(3.141592654d0 * (i - 0.25d0) / (n + 0.5d0))
       pp = 0.d0
       do while (abs(z - z1) > 3.d-14)
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
       end do
       x(i) = xm - xl * z
       x(n + 1 - i) = xm + xl * z
       w(i) = 2.d0 * xl / ((1.d0 - z * z) * pp * pp)
       w(n + 1 - i) = w(i)
    end do
  end subroutine gauleg

end program ccm_gpu
