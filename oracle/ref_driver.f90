! ref_driver.f90 -- driver for the reference Fortran kMC modules.
!
! TEST INFRASTRUCTURE ONLY.  Linked against the checked-in generated
! Fortran of a kmos model (compiled in place by oracle/build_ref.py with the
! injected uniform of oracle/ref_rng.c) to produce reference trajectories.
!
! Input (list-directed, one file):
!   X Y Z
!   seed stream
!   nsteps mode            mode 0: per-step trace, 1: one do_kmc_steps call
!                          (its wall time is printed on stdout), 2: trace,
!                          then save_system (writes ref.reload in the
!                          working directory), 3: deallocate_system and
!                          reload_system('in') (reads in.reload), then trace
!   has_config
!   rates(1:nr_of_proc)
!   config(1:volume)       species per site, only if has_config /= 0
!   rng_step               mode 3 only: first step of the uniform stream
! Output (stream access, little endian), see tests/refbin.py:
!   int32 nexec, [int32 proc(n), int32 site(n), real64 time(n)] (mode 0)
!   real64 kmc_time, int64 kmc_step, int32 lattice(volume),
!   int32 nr_of_sites(P), int64 procstat(P), real64 integ_rates(P),
!   int32 avail_sites(P, volume, 2)
!
! The trace loop issues the same calls, in the same order, as
! do_kmc_step (kmos/fortran_src/proclist_generic_subroutines.mpy:47-80);
! mode 1 runs the reference's own do_kmc_steps so both can be compared.
program ref_driver
  use iso_c_binding, only: c_int32_t, c_int64_t
  use kind_values
  use base, only: set_rate_const, update_accum_rate, update_clocks, &
       update_integ_rate, determine_procsite, get_kmc_time, get_kmc_step, &
       get_nrofsites, get_procstat, get_integ_rate, avail_sites, &
       base_get_species => get_species, base_replace_species => replace_species, &
       get_volume, deallocate_system, reload_system, save_system
  use proclist, only: init, run_proc_nr, do_kmc_steps, nr_of_proc
  use lattice, only: default_layer
  implicit none
  interface
    subroutine kmcg_ref_rng_set(seed, stream) bind(C, name="kmcg_ref_rng_set")
      import :: c_int32_t
      integer(c_int32_t), value :: seed, stream
    end subroutine
    subroutine kmcg_ref_rng_seek(step) bind(C, name="kmcg_ref_rng_seek")
      import :: c_int64_t
      integer(c_int64_t), value :: step
    end subroutine
  end interface
  character(len=4096) :: inpath, outpath
  character(len=200) :: sysname, loadname
  logical :: reloaded
  integer(kind=ilong) :: rng_step
  integer(kind=iint) :: nx, ny, nz, seed, stream, mode, has_config
  integer(kind=ilong) :: nsteps, i, kstep
  integer(kind=iint) :: p, s, volume, proc_nr, nr_site, sp, x, y, z, n
  integer(kind=iint), dimension(2) :: size2
  real(kind=rdouble), allocatable :: rates(:), times(:)
  integer(kind=iint), allocatable :: config(:), tproc(:), tsite(:)
  real(kind=rdouble) :: ran_time, ran_proc, ran_site, t, r
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

  sysname = 'ref'
  size2 = (/ nx, ny /)
  call kmcg_ref_rng_set(int(seed, c_int32_t), int(stream, c_int32_t))
  call init(size2, sysname, default_layer, seed, .true.)
  call get_volume(volume)
  do p = 1, nr_of_proc
    call set_rate_const(p, rates(p))
  end do
  if (has_config /= 0) then
    allocate(config(volume))
    read(u_in, *) config
    do s = 1, volume
      sp = base_get_species(s)
      call base_replace_species(s, sp, config(s))
    end do
    ! _adjust_database: x outermost (kmos/run/__init__.py:1439-1457)
    do x = 0, nx - 1
      do y = 0, ny - 1
        do z = 0, nz - 1
          call ref_touchup(x, y, z)
        end do
      end do
    end do
  end if
  rng_step = 0
  if (mode == 3) then
    read(u_in, *) rng_step
    call deallocate_system()
    loadname = 'in'
    call reload_system(loadname, reloaded)
    if (.not. reloaded) stop 'ref_driver: in.reload not found'
  end if
  close(u_in)
  ! drawing starts after set-up, as in the engine
  call kmcg_ref_rng_set(int(seed, c_int32_t), int(stream, c_int32_t))
  call kmcg_ref_rng_seek(int(rng_step, c_int64_t))

  open(newunit=u_out, file=trim(outpath), access='stream', form='unformatted', &
       status='replace', action='write')
  if (mode /= 1) then
    allocate(tproc(nsteps), tsite(nsteps), times(nsteps))
    do i = 1, nsteps
      call kmcg_ref_uniform(ran_time)
      call kmcg_ref_uniform(ran_proc)
      call kmcg_ref_uniform(ran_site)
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
    if (mode == 2) call save_system()
  else
    call system_clock(c0, crate)
    call do_kmc_steps(nsteps)
    call system_clock(c1)
    ! wall seconds of the reference's own step loop (the CPU baseline)
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
  close(u_out)
end program ref_driver
