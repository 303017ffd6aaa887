! ref_driver_otf.f90 -- driver for the reference's generated otf Fortran.
!
! TEST INFRASTRUCTURE ONLY.  Linked against the checked-in generated otf
! Fortran of a kmos model (compiled in place by oracle/build_ref.py with
! the injected uniform of oracle/ref_rng.c) to produce
! reference trajectories.
!
! Input (list-directed, one file):
!   X Y Z
!   seed stream
!   nsteps mode            mode 0: per-step trace, 1: one do_kmc_steps call
!                          (its wall time is printed on stdout)
!   has_config
!   rates(1:nr_of_proc)
!   npar
!   userpar(1:npar)        only if npar > 0 (sorted parameter names)
!   config(1:volume)       species per site, only if has_config /= 0
! Set-up: init, set_rate_const and update_user_parameter for every entry,
! recalculate_rates_matrix (what kmos.run's set_rate_constants does for
! otf, kmos/run/__init__.py:2415-2432), then the configuration through
! replace_species and touchup_cell with x outermost (_adjust_database,
! kmos/run/__init__.py:1439-1457).
! Output (stream access, little endian), see tests/refbin.py:
!   int32 nexec, [int32 proc(n), int32 site(n), real64 time(n)] (mode 0)
!   real64 kmc_time, int64 kmc_step, int32 lattice(volume),
!   int32 nr_of_sites(P), int64 procstat(P), real64 integ_rates(P),
!   int32 avail_sites(P, volume, 2), real64 rates_matrix(p, 1:volume+1)
!   for p = 1..P (get_rate)
!
! The trace loop issues the same calls, in the same order, as the
! fixtures' do_kmc_step (proclist.f90).
program ref_driver_otf
  use iso_c_binding, only: c_int32_t, c_int64_t
  use kind_values
  use base, only: set_rate_const, update_accum_rate, update_clocks, &
       update_integ_rate, determine_procsite, get_kmc_time, get_kmc_step, &
       get_nrofsites, get_procstat, get_integ_rate, avail_sites, get_rate, &
       base_get_species => get_species, base_replace_species => replace_species, &
       get_volume
  use proclist, only: init, run_proc_nr, do_kmc_steps, nr_of_proc, touchup_cell, &
       recalculate_rates_matrix
  use proclist_pars, only: update_user_parameter
  use lattice, only: default_layer
  implicit none
  interface
    subroutine kmcg_ref_rng_set(seed, stream) bind(C, name="kmcg_ref_rng_set")
      import :: c_int32_t
      integer(c_int32_t), value :: seed, stream
    end subroutine
  end interface
  character(len=4096) :: inpath, outpath
  character(len=400) :: sysname
  integer(kind=iint) :: nx, ny, nz, seed, stream, mode, has_config, npar
  integer(kind=ilong) :: nsteps, i, kstep
  integer(kind=iint) :: p, s, volume, proc_nr, nr_site, sp, x, y, z, k
  integer(kind=iint), dimension(2) :: size2
  real(kind=rdouble), allocatable :: rates(:), times(:), userpar(:)
  integer(kind=iint), allocatable :: config(:), tproc(:), tsite(:)
  real(kind=rsingle) :: ran_time, ran_proc, ran_site
  real(kind=rdouble) :: t, r
  integer(kind=ilong) :: ps
  integer(kind=iint) :: nrs, u_in, u_out, nexec
  integer(kind=8) :: c0, c1, crate

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  open(newunit=u_in, file=trim(inpath), status='old', action='read')
  read(u_in, *) nx, ny, nz
  read(u_in, *) seed, stream
  read(u_in, *) nsteps, mode
  read(u_in, *) has_config
  allocate(rates(nr_of_proc))
  read(u_in, *) rates
  read(u_in, *) npar
  allocate(userpar(max(npar, 1)))
  if (npar > 0) read(u_in, *) userpar(1:npar)

  sysname = 'ref'
  size2 = (/ nx, ny /)
  call init(size2, sysname, default_layer, seed, .true.)
  call get_volume(volume)
  do p = 1, nr_of_proc
    call set_rate_const(p, rates(p))
  end do
  do k = 1, npar
    call update_user_parameter(k, userpar(k))
  end do
  call recalculate_rates_matrix()
  if (has_config /= 0) then
    allocate(config(volume))
    read(u_in, *) config
    do s = 1, volume
      sp = base_get_species(s)
      call base_replace_species(s, sp, config(s))
    end do
    do x = 0, nx - 1
      do y = 0, ny - 1
        do z = 0, nz - 1
          call touchup_cell((/ x, y, z, 0 /))
        end do
      end do
    end do
  end if
  close(u_in)
  ! drawing starts after set-up, as in the engine
  call kmcg_ref_rng_set(int(seed, c_int32_t), int(stream, c_int32_t))

  open(newunit=u_out, file=trim(outpath), access='stream', form='unformatted', &
       status='replace', action='write')
  if (mode /= 1) then
    allocate(tproc(nsteps), tsite(nsteps), times(nsteps))
    do i = 1, nsteps
      call random_number(ran_time)
      call random_number(ran_proc)
      call random_number(ran_site)
      call update_accum_rate
      call update_clocks(ran_time)
      call update_integ_rate
      call determine_procsite(ran_proc, ran_site, proc_nr, nr_site)
      call run_proc_nr(proc_nr, nr_site)
      tproc(i) = proc_nr
      tsite(i) = nr_site
      call get_kmc_time(t)
      times(i) = t
    end do
    nexec = int(nsteps, iint)
    write(u_out) nexec
    write(u_out) tproc, tsite, times
  else
    call system_clock(c0, crate)
    call do_kmc_steps(nsteps)
    call system_clock(c1)
    write(*, '(A, ES24.16)') 'do_kmc_steps_seconds ', real(c1 - c0, rdouble) / real(crate, rdouble)
    nexec = int(nsteps, iint)
    write(u_out) nexec
  end if

  call get_kmc_time(t)
  call get_kmc_step(kstep)
  write(u_out) t, kstep
  do s = 1, volume
    write(u_out) base_get_species(s)
  end do
  do p = 1, nr_of_proc
    call get_nrofsites(p, nrs)
    write(u_out) nrs
  end do
  do p = 1, nr_of_proc
    call get_procstat(p, ps)
    write(u_out) ps
  end do
  do p = 1, nr_of_proc
    call get_integ_rate(p, r)
    write(u_out) r
  end do
  write(u_out) avail_sites
  do p = 1, nr_of_proc
    do s = 1, volume + 1
      if (s <= volume) then
        call get_rate(p, s, r)
      else
        call get_rate(p, -1, r)
      end if
      write(u_out) r
    end do
  end do
  close(u_out)
end program ref_driver_otf
