! A Fortran caller of libnb6gpu.so, shaped like the integrator's own calls:
!   the regular block of CALC_REG_ON_GPU (util_gpu.F:25-111): GPUNB_REGF on chunks
!   of up to MAXTHR particles with LISTGP(LMAX,MAXTHR) in column-major order, the
!   overflow rescale of RS and re-run of the chunk (util_gpu.F:65-100), +IFIRST and
!   self removal (util_gpu.F:101-110), PHIDBL from the GPU potential;
!   the energy check's GPUPOT (energy.F:33);
!   the irregular block's IRR_SIMD_SET_JP / SET_LIST / FIRR_VEC (intgrt.F:533,932).
! Every argument is passed by reference, as the reference's Fortran does.
! Usage: regf_caller <input file> <output file> (unformatted stream, native endian).
program regf_caller
  implicit none
  integer, parameter :: MAXTHR = 1024
  integer :: n, nnbopt, nnbmax, lmax, ifirst, nireg, mflag, nbflag, nirr, irank
  integer :: nretry, i, k, u, nchunk, istart
  double precision :: bodym, ti
  double precision, allocatable :: body(:), x(:,:), xdot(:,:), rs(:), stepr(:)
  double precision, allocatable :: f(:,:), fdot(:,:), t0(:), phidbl(:), pot(:)
  double precision, allocatable :: acc(:,:), jrk(:,:), ai(:,:), ji(:,:)
  integer, allocatable :: ireg(:), iaddr(:), lists(:,:), nnbid(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, nnbopt, nnbmax, lmax, ifirst, nireg, mflag, nbflag, nirr
  allocate(body(n), x(3,n), xdot(3,n), rs(n), stepr(n), f(3,n), fdot(3,n), t0(n))
  allocate(phidbl(n), pot(n), ireg(nireg), iaddr(nirr), lists(lmax,nireg))
  allocate(acc(3,nireg), jrk(3,nireg), ai(3,nirr), ji(3,nirr), nnbid(nirr))
  read(u) body, x, xdot, rs, stepr, bodym, f, fdot, t0, ti
  read(u) ireg, iaddr
  close(u)

  irank = 0
  call gpunb_open(n, irank)
  call gpunb_send(n, body, x, xdot)
  nretry = 0
  phidbl = 0.0d0
  do istart = 1, nireg, MAXTHR
    nchunk = min(MAXTHR, nireg - istart + 1)
    call reg_chunk(ireg(istart:istart+nchunk-1), nchunk)
  end do
  call gpunb_profile(irank)
  call gpunb_close()

  call gpupot(irank, 1, n, n, body, x, pot)

  call irr_simd_open(n, lmax, irank)
  do i = 1, n
    call irr_simd_set_jp(i, x(1,i), xdot(1,i), f(1,i), fdot(1,i), body(i), t0(i))
  end do
  do k = 1, nireg
    call irr_simd_set_list(ireg(k), lists(1,k))
  end do
  call irr_simd_firr_vec(ti, nirr, iaddr, ai, ji, nnbid)
  call irr_simd_profile(irank)
  call irr_simd_close(irank)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) nretry
  write(u) acc, jrk, phidbl, rs, pot, ai, ji
  write(u) lists, nnbid
  close(u)

contains

  ! One chunk of the regular block: the i-particles idx(1:ni).
  subroutine reg_chunk(idx, ni)
    integer, intent(in) :: ni
    integer, intent(in) :: idx(ni)
    integer :: listgp(lmax, MAXTHR)
    double precision :: h2(MAXTHR), dtr(MAXTHR), xi(3,MAXTHR), vi(3,MAXTHR)
    double precision :: ga(3,MAXTHR), gj(3,MAXTHR), gp(MAXTHR)
    integer :: ii, id, nnb, ll, l1, jt
    logical :: again
    do ii = 1, ni
      id = idx(ii)
      if (mflag == 1) then
        h2(ii) = rs(id)**2 / bodym
      else
        h2(ii) = rs(id)**2
      end if
      dtr(ii) = stepr(id)
      xi(:,ii) = x(:,id)
      vi(:,ii) = xdot(:,id)
    end do
    again = .true.
    do while (again)
      call gpunb_regf(ni, h2, dtr, xi, vi, ga, gj, gp, lmax, nnbmax, listgp, mflag)
      again = .false.
      ! the first overflowing row rescales its RS and re-runs the whole chunk
      do ii = 1, ni
        nnb = listgp(1, ii)
        if (nnb < 0) then
          id = idx(ii)
          if (nbflag == 1) then
            if (-nnb > nnbopt) then
              rs(id) = (-real(nnbopt) / real(nnb))**0.333 * rs(id)
            else
              rs(id) = (real(nnbopt) / real(nnbmax))**0.4 * rs(id)
            end if
          else
            rs(id) = 0.9 * rs(id)
          end if
          if (mflag == 1) then
            h2(ii) = rs(id)**2 / bodym
          else
            h2(ii) = rs(id)**2
          end if
          nretry = nretry + 1
          again = .true.
          exit
        end if
      end do
    end do
    do ii = 1, ni
      id = idx(ii)
      nnb = listgp(1, ii)
      l1 = 1
      do ll = 2, nnb + 1
        jt = listgp(ll, ii) + ifirst     ! GPU indices start at 0
        if (jt /= id) then
          l1 = l1 + 1
          listgp(l1, ii) = jt
        end if
      end do
      listgp(1, ii) = l1 - 1
      phidbl(id) = gp(ii)
      acc(:, istart + ii - 1) = ga(:, ii)
      jrk(:, istart + ii - 1) = gj(:, ii)
      lists(:, istart + ii - 1) = listgp(:, ii)
    end do
  end subroutine reg_chunk

end program regf_caller
