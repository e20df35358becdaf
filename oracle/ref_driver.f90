! Oracle harness (test infrastructure, never shipped or measured).
!
! Drives the compiled reference code (object files built by oracle/Makefile from
! the reference sources in place) to produce golden vectors:
!   * the reference's own setup sequence of the PBC branch of ccm_main.f90:281-317,
!   * the CCD loop of ccd_iter (ccm_iter.f90:1-170) with the DIIS subspace/cadence,
!     tolerance and iteration cap exposed as arguments (the reference hard-codes
!     6/6, 1e-6 and 100 at ccm_iter.f90:44-45, 71-74),
!   * an optional twist vector (the reference reads twist_angle(:,1) in
!     setup_sp_data, ccm_main.f90:1036-1044), or the twist of one TABC
!     combination: the reference's gauss_legendre nodes on [0, pi]
!     (ccm_main.f90:909-921) and its (nx <= ny <= nz) enumeration
!     (ccm_main.f90:169-204), with the weight factor w w w / pi^3 (236-242),
!   * an optional swap of every NN block for the reference's own Minnesota
!     function vmom_minnesota (chiral_module_andreas_with_delta.f90:8980-9111),
!   * the (T) energy through pert_triples2 (ccm_triples.f90:528-735),
!   * binary dumps of orbitals, Fock matrix, canonical V elements and converged T2.
!
! Usage: ref_driver ccm_in outdir interaction subspace diis_step tol maxit
!                   twx twy twz dump triples [ntwist combination]
subroutine ref_driver_main() bind(C, name='ref_driver_main')
  use parallel
  use constants
  use single_particle_orbits
  use one_body_operators
  use t2_storage
  use diis_mod
  use chiral_constants
  use chiral_potentials
  use ang_mom_functions
  use wave_functions, only: gauss_legendre
  implicit none
  character(len=512) :: inputfile, outdir, arg, interaction
  integer :: nsub, nstepdiis, maxit, dodump, dotriples, count, itimes, u
  real*4 :: tol4
  real*8 :: tstart, tloop
  real*8 :: tol, tw(3), tabc_w
  integer :: ntw, combo, ia, ib, ic, kc
  real*8, allocatable :: xg(:), wg(:)
  complex*16 :: ener1, ener2, dener

  call mpi_init(ierror)
  call mpi_comm_size(mpi_comm_world, num_procs, ierror)
  call mpi_comm_rank(mpi_comm_world, iam, ierror)
  chiral_delta_flag = .true.
  master = 0

  call get_command_argument(1, inputfile)
  call get_command_argument(2, outdir)
  call get_command_argument(3, interaction)
  call get_command_argument(4, arg); read(arg,*) nsub
  call get_command_argument(5, arg); read(arg,*) nstepdiis
  call get_command_argument(6, arg); read(arg,*) tol4
  call get_command_argument(7, arg); read(arg,*) maxit
  call get_command_argument(8, arg); read(arg,*) tw(1)
  call get_command_argument(9, arg); read(arg,*) tw(2)
  call get_command_argument(10, arg); read(arg,*) tw(3)
  call get_command_argument(11, arg); read(arg,*) dodump
  call get_command_argument(12, arg); read(arg,*) dotriples
  ntw = 0
  combo = 0
  tabc_w = 1.d0
  if (command_argument_count() >= 14) then
     call get_command_argument(13, arg); read(arg,*) ntw
     call get_command_argument(14, arg); read(arg,*) combo
  end if
  if (ntw > 0) then
     allocate(xg(ntw), wg(ntw))
     xg = 0.d0; wg = 0.d0
     call gauss_legendre(0.d0, pi, xg, wg, ntw)
     kc = 0
     do ia = 1, ntw
        do ib = ia, ntw
           do ic = ib, ntw
              kc = kc + 1
              if (kc == combo) then
                 tw = (/ xg(ia), xg(ib), xg(ic) /)
                 tabc_w = 6.d0
                 if (ia == ib .or. ib == ic) tabc_w = 3.d0
                 if (ia == ib .and. ib == ic) tabc_w = 1.d0
                 tabc_w = tabc_w * wg(ia) * wg(ib) * wg(ic) / (1.d0 * pi)**3
              end if
           end do
        end do
     end do
  end if
  ! the reference compares |dE| against a default-real literal (ccm_iter.f90:72)
  tol = real(tol4, 8)

  open(5, file=trim(inputfile), access="SEQUENTIAL")
  read(5,*); read(5,*) delta_chiral_order, lambda_delta
  read(5,*); read(5,*) cE, cD
  if (delta_chiral_order < NNLO) then
     cE = 0.d0
     cD = 0.d0
  end if
  call commons_to_angmom
  call init_chp_constants
  call allocate_sp_data

  call setup_N3LO_int_mesh(10)
  twist_angle = 0.d0
  twist_angle(1,1) = tw(1)
  twist_angle(2,1) = tw(2)
  twist_angle(3,1) = tw(3)
  call setup_sp_data(1,1,1)
  call precalc_chp_functions
  call setup_channel_structures
  call setup_ph_channel_structures
  if (trim(interaction) == 'minnesota') call swap_in_minnesota(.false.)
  call normal_ordered_hamiltonian
  call setup_t_amplitudes

  ! only rank 0 writes; the reference routines run SPMD on every rank
  if (iam /= 0) then
     open(newunit=u, file='/dev/null', status='old')
  else
     open(newunit=u, file=trim(outdir)//'/energies.txt', status='replace')
  end if
  write(u,'(a,2es26.17)') 'e0 ', real(e0), aimag(e0)
  write(u,'(a,4es26.17)') 'twist ', tw(1), tw(2), tw(3), tabc_w
  call ccd_energy_save(ener1)
  mbpt2 = e0 + ener1
  write(u,'(a,i6,2es26.17)') 'it ', 0, real(ener1), aimag(ener1)

  if (dodump /= 0 .and. iam == 0) call dump_setup(trim(outdir))

  subspace = nsub
  diis_step = nstepdiis
  call diis_setup
  allocate( d1ik(1:below_ef), d1ac(below_ef+1:tot_orbs) )
  allocate( t1_denom(below_ef+1:tot_orbs, 1:below_ef) )
  allocate( d2ij(1:below_ef), d2ab(below_ef+1:tot_orbs) )
  t1_denom = 0.d0; d1ik = 0.d0; d1ac = 0.d0; d2ij = 0.d0; d2ab = 0.d0
  nstep = 0
  count = 0
  ener2 = 0.d0
  dener = 4.0
  call mpi_barrier(mpi_comm_world, ierror)
  tstart = MPI_WTIME()
  do itimes = 1, maxit
     if (abs(dener) > tol) then
        call t2_intermediate(0)
        count = count + 1
        nstep = nstep + 1
        call diis_t1t2(count)
        call ccd_energy_save(ener2)
        dener = ener2 - ener1
        ener1 = ener2
        write(u,'(a,i6,2es26.17)') 'it ', itimes, real(ener2), aimag(ener2)
     end if
  end do
  call mpi_barrier(mpi_comm_world, ierror)
  tloop = MPI_WTIME() - tstart
  write(u,'(a,es26.17,i6)') 'looptime ', tloop, num_procs
  call diis_take_down
  eccsd = e0 + ener2
  write(u,'(a,2es26.17)') 'eccd ', real(eccsd), aimag(eccsd)
  write(u,'(a,i6)') 'a ', below_ef
  write(u,'(a,2es26.17)') 'kf ', kf, dens
  write(u,'(a,2es26.17)') 'box ', lx, volume

  if (dodump /= 0 .and. iam == 0) call dump_t2(trim(outdir))

  if (dotriples /= 0) then
     call setup_vnn_hppp_block
     if (trim(interaction) == 'minnesota') call swap_in_minnesota(.true.)
     if (dodump /= 0 .and. iam == 0) call dump_hppp(trim(outdir))
     call setup_t3_channel_structures
     call pert_triples2
     write(u,'(a,2es26.17)') 'eccdt ', real(eccsdt), aimag(eccsdt)
  end if
  close(u)
  call mpi_finalize(ierror)

contains

  complex*16 function vmin(p, q, r, s)
    integer, intent(in) :: p, q, r, s
    vmin = dcmplx(vmom_minnesota(p,q,r,s) - vmom_minnesota(p,q,s,r), 0.d0)
  end function vmin

  ! Overwrite every stored NN block with the antisymmetrized reference Minnesota
  ! element, using the reference's own lookup tables.
  subroutine swap_in_minnesota(hppp_only)
    logical, intent(in) :: hppp_only
    integer :: ch, bra, ket, p, q, r, s
    if (hppp_only) then
       do ch = 1, size(vnn_hppp)
          do ket = 1, size(lookup_hppp_configs(2,ch)%ival, 2)
             p = lookup_hppp_configs(2,ch)%ival(1,ket); q = lookup_hppp_configs(2,ch)%ival(2,ket)
             do bra = 1, size(lookup_hppp_configs(1,ch)%ival, 2)
                r = lookup_hppp_configs(1,ch)%ival(1,bra); s = lookup_hppp_configs(1,ch)%ival(2,bra)
                vnn_hppp(ch)%val(bra,ket) = vmin(p,q,r,s)
             end do
          end do
       end do
       return
    end if
    do ch = 1, size(vnn_hhhh)
       do bra = 1, size(lookup_hhhh_configs(1,ch)%ival, 2)
          p = lookup_hhhh_configs(1,ch)%ival(1,bra); q = lookup_hhhh_configs(1,ch)%ival(2,bra)
          do ket = 1, size(lookup_hhhh_configs(1,ch)%ival, 2)
             r = lookup_hhhh_configs(1,ch)%ival(1,ket); s = lookup_hhhh_configs(1,ch)%ival(2,ket)
             vnn_hhhh(ch)%val(bra,ket) = vmin(p,q,r,s)
          end do
       end do
    end do
    do ch = 1, size(vnn_hhhp)
       do bra = 1, size(lookup_hhhp_configs(1,ch)%ival, 2)
          p = lookup_hhhp_configs(1,ch)%ival(1,bra); q = lookup_hhhp_configs(1,ch)%ival(2,bra)
          do ket = 1, size(lookup_hhhp_configs(2,ch)%ival, 2)
             r = lookup_hhhp_configs(2,ch)%ival(1,ket); s = lookup_hhhp_configs(2,ch)%ival(2,ket)
             vnn_hhhp(ch)%val(bra,ket) = vmin(p,q,r,s)
          end do
       end do
    end do
    do ch = 1, size(vnn_hhpp)
       do bra = 1, size(lookup_hhpp_configs(1,ch)%ival, 2)
          p = lookup_hhpp_configs(1,ch)%ival(1,bra); q = lookup_hhpp_configs(1,ch)%ival(2,bra)
          do ket = 1, size(lookup_hhpp_configs(2,ch)%ival, 2)
             r = lookup_hhpp_configs(2,ch)%ival(1,ket); s = lookup_hhpp_configs(2,ch)%ival(2,ket)
             vnn_hhpp(ch)%val(bra,ket) = vmin(p,q,r,s)
          end do
       end do
    end do
    do ch = 1, size(vnn_hphp)
       do bra = 1, size(lookup_hphp_configs(1,ch)%ival, 2)
          p = lookup_hphp_configs(1,ch)%ival(1,bra); q = lookup_hphp_configs(1,ch)%ival(2,bra)
          do ket = 1, size(lookup_hphp_configs(1,ch)%ival, 2)
             r = lookup_hphp_configs(1,ch)%ival(1,ket); s = lookup_hphp_configs(1,ch)%ival(2,ket)
             vnn_hphp(ch)%val(bra,ket) = vmin(p,q,r,s)
          end do
       end do
    end do
    do ch = 1, size(lookup_pppp_configs, 2)
       if (.not. allocated(vnn_pppp(ch)%val)) cycle
       do bra = lbound(vnn_pppp(ch)%val,1), ubound(vnn_pppp(ch)%val,1)
          p = lookup_pppp_configs(1,ch)%ival(1,bra); q = lookup_pppp_configs(1,ch)%ival(2,bra)
          do ket = lbound(vnn_pppp(ch)%val,2), ubound(vnn_pppp(ch)%val,2)
             r = lookup_pppp_configs(1,ch)%ival(1,ket); s = lookup_pppp_configs(1,ch)%ival(2,ket)
             vnn_pppp(ch)%val(bra,ket) = vmin(p,q,r,s)
          end do
       end do
    end do
  end subroutine swap_in_minnesota

  ! Orbitals, Fock matrix and the canonical (first index < second index) V
  ! elements of the hhhh, hhpp, hphp and pppp blocks, as little-endian streams.
  subroutine dump_setup(d)
    character(len=*), intent(in) :: d
    integer :: v, ch, bra, ket, n
    integer :: p, q, r, s
    open(newunit=v, file=d//'/orbitals.bin', access='stream', form='unformatted', status='replace')
    write(v) tot_orbs, below_ef
    write(v) all_orbit%nx(1:tot_orbs), all_orbit%ny(1:tot_orbs), all_orbit%nz(1:tot_orbs)
    write(v) all_orbit%szp(1:tot_orbs), all_orbit%itzp(1:tot_orbs)
    write(v) all_orbit%kx(1:tot_orbs), all_orbit%ky(1:tot_orbs), all_orbit%kz(1:tot_orbs), all_orbit%e(1:tot_orbs)
    write(v) lx, volume, kf, dens
    close(v)
    open(newunit=v, file=d//'/fock.bin', access='stream', form='unformatted', status='replace')
    write(v) fock_mtx
    write(v) e0
    close(v)
    open(newunit=v, file=d//'/vblocks.bin', access='stream', form='unformatted', status='replace')
    ! kind 1 = hhhh, 2 = hhhp, 3 = hhpp, 4 = hphp, 6 = pppp; records (kind,p,q,r,s,re,im)
    n = 0
    do ch = 1, size(vnn_hhhh)
       do bra = 1, size(lookup_hhhh_configs(1,ch)%ival, 2)
          p = lookup_hhhh_configs(1,ch)%ival(1,bra); q = lookup_hhhh_configs(1,ch)%ival(2,bra)
          if (p >= q) cycle
          do ket = 1, size(lookup_hhhh_configs(1,ch)%ival, 2)
             r = lookup_hhhh_configs(1,ch)%ival(1,ket); s = lookup_hhhh_configs(1,ch)%ival(2,ket)
             if (r >= s) cycle
             write(v) 1, p, q, r, s, vnn_hhhh(ch)%val(bra,ket)
          end do
       end do
    end do
    do ch = 1, size(vnn_hhpp)
       do bra = 1, size(lookup_hhpp_configs(1,ch)%ival, 2)
          p = lookup_hhpp_configs(1,ch)%ival(1,bra); q = lookup_hhpp_configs(1,ch)%ival(2,bra)
          if (p >= q) cycle
          do ket = 1, size(lookup_hhpp_configs(2,ch)%ival, 2)
             r = lookup_hhpp_configs(2,ch)%ival(1,ket); s = lookup_hhpp_configs(2,ch)%ival(2,ket)
             if (r >= s) cycle
             write(v) 3, p, q, r, s, vnn_hhpp(ch)%val(bra,ket)
          end do
       end do
    end do
    do ch = 1, size(vnn_hhhp)
       do bra = 1, size(lookup_hhhp_configs(1,ch)%ival, 2)
          p = lookup_hhhp_configs(1,ch)%ival(1,bra); q = lookup_hhhp_configs(1,ch)%ival(2,bra)
          if (p >= q) cycle
          do ket = 1, size(lookup_hhhp_configs(2,ch)%ival, 2)
             r = lookup_hhhp_configs(2,ch)%ival(1,ket); s = lookup_hhhp_configs(2,ch)%ival(2,ket)
             write(v) 2, p, q, r, s, vnn_hhhp(ch)%val(bra,ket)
          end do
       end do
    end do
    do ch = 1, size(vnn_hphp)
       do bra = 1, size(lookup_hphp_configs(1,ch)%ival, 2)
          p = lookup_hphp_configs(1,ch)%ival(1,bra); q = lookup_hphp_configs(1,ch)%ival(2,bra)
          do ket = 1, size(lookup_hphp_configs(1,ch)%ival, 2)
             r = lookup_hphp_configs(1,ch)%ival(1,ket); s = lookup_hphp_configs(1,ch)%ival(2,ket)
             write(v) 4, p, q, r, s, vnn_hphp(ch)%val(bra,ket)
          end do
       end do
    end do
    do ch = 1, size(lookup_pppp_configs, 2)
       if (.not. allocated(vnn_pppp(ch)%val)) cycle
       do bra = lbound(vnn_pppp(ch)%val,1), ubound(vnn_pppp(ch)%val,1)
          p = lookup_pppp_configs(1,ch)%ival(1,bra); q = lookup_pppp_configs(1,ch)%ival(2,bra)
          if (p >= q) cycle
          do ket = lbound(vnn_pppp(ch)%val,2), ubound(vnn_pppp(ch)%val,2)
             r = lookup_pppp_configs(1,ch)%ival(1,ket); s = lookup_pppp_configs(1,ch)%ival(2,ket)
             if (r >= s) cycle
             write(v) 6, p, q, r, s, vnn_pppp(ch)%val(bra,ket)
          end do
       end do
    end do
    close(v)
  end subroutine dump_setup

  ! Canonical (a<b, i<j) converged T2 elements as records (a,b,i,j,re,im).
  subroutine dump_t2(d)
    character(len=*), intent(in) :: d
    integer :: v, ch, bra, ket, a, b, i, j
    open(newunit=v, file=d//'/t2.bin', access='stream', form='unformatted', status='replace')
    do ch = 1, size(t2_ccm)
       do bra = 1, size(lookup_hhpp_configs(2,ch)%ival, 2)
          a = lookup_hhpp_configs(2,ch)%ival(1,bra); b = lookup_hhpp_configs(2,ch)%ival(2,bra)
          if (a >= b) cycle
          do ket = 1, size(lookup_hhpp_configs(1,ch)%ival, 2)
             i = lookup_hhpp_configs(1,ch)%ival(1,ket); j = lookup_hhpp_configs(1,ch)%ival(2,ket)
             if (i >= j) cycle
             write(v) a, b, i, j, t2_ccm(ch)%val(bra,ket)
          end do
       end do
    end do
    close(v)
  end subroutine dump_t2

  ! hppp elements used by (T): records (k,d,c,b,re,im) with value <cb||kd>.
  subroutine dump_hppp(d)
    character(len=*), intent(in) :: d
    integer :: v, ch, bra, ket, p, q, r, s
    open(newunit=v, file=d//'/vhppp.bin', access='stream', form='unformatted', status='replace')
    do ch = 1, size(vnn_hppp)
       do ket = 1, size(lookup_hppp_configs(2,ch)%ival, 2)
          p = lookup_hppp_configs(2,ch)%ival(1,ket); q = lookup_hppp_configs(2,ch)%ival(2,ket)
          if (p >= q) cycle
          do bra = 1, size(lookup_hppp_configs(1,ch)%ival, 2)
             r = lookup_hppp_configs(1,ch)%ival(1,bra); s = lookup_hppp_configs(1,ch)%ival(2,bra)
             write(v) r, s, p, q, vnn_hppp(ch)%val(bra,ket)
          end do
       end do
    end do
    close(v)
  end subroutine dump_hppp

end subroutine ref_driver_main
