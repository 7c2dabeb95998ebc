      subroutine vmc
c Drop-in replacement of the reference's vmc (reptation.f:973-1014)
c whose steps run on the GPU.  Same block structure: read_conf and
c getconf, averages(1) and averages(3) around every block, putconf
c every nskip steps and at the end when the stack is empty, write_conf.
c The steps themselves (move, metropolis_test, averages(2);
c reptation.f:993-1002) are qmc_vmc_sweep calls on one walker: the
c walker (x_old) and its random stream (/rnyucm/) go to the device
c once, qmc_vmc_start evaluates it, each block is one sweep of nstp
c steps (split where putconf needs the walker every nskip steps), and
c the sweep's per-step sums of p_old are folded into averages(2) as
c nstp steps of weight 1 (the block's sum once, then nstp - 1 zero
c steps, so that blk_av and blk_norm hold exactly what the reference's
c per-step calls leave).  The walker comes back to COMMON (x_old,
c g_old, h_old, p_old, age, grad2_ph and the stream) where putconf or
c the caller reads it.  Link it, with compute_properties_hip.f whose
c device context it uses, instead of the reference's vmc.
c
c compute_vmcder (der_nskip > 0) is not on the device: such a deck runs
c the reference's step loop (vmc_cpu, the reference's vmc renamed),
c whose evaluations are on the GPU through compute_properties_hip.f.
c bm_order 0 draws the Box-Muller uniforms in the order the amdflang
c build of the reference does.
      use iso_c_binding
      use qmc_hip_mod
      use qmc_hip_host
      implicit none
      include 'ewald.h'
      integer iblk, istp, nitc0, ncmass0, nmstar0, i, rc, nchunk
      integer mrn(4), lrn(4), nbitrn, irnyrn
      common /rnyucm/ mrn, lrn, nbitrn, irnyrn
      real*8 uno, psave(m_props)
      real(c_double), allocatable, target :: hx(:,:), hg(:,:), hp(:)
      real(c_double), allocatable, target :: hb(:), hz(:)
      real(c_double), target :: hh(2), hg2(1), hage(1)
      integer(c_int64_t), target :: hrng(1)
      type(c_ptr) :: d_x, d_g, d_h, d_p, d_g2, d_age, d_rng, d_blk
      type(qmc_walkers) :: w
      integer(c_size_t) :: b2n, bnp
      if (der_nskip .ne. 0) then
       call vmc_cpu
       return
      endif
      uno = 1.d0
      if (ncmass .ne. 0) then
       ncmass0 = ncmass
       ncmass = 0
      endif
      if (nmstar .ne. 0) then
       nmstar0 = nmstar
       nmstar = 0
      endif
      if (nitc .ne. 0) then
       nitc0 = nitc
       nitc = 0
      endif
      call read_conf
      call getconf
      if (.not. c_associated(ctx)) stop 'vmc (GPU): no device context'
      allocate(hx(2,nptot), hg(2,nptot), hp(info%n_props),
     &         hb(info%n_props), hz(info%n_props))
      hz = 0.d0
      b2n = 8_c_size_t * 2 * nptot
      bnp = 8_c_size_t * info%n_props
      rc = hipMalloc(d_x, b2n)
      rc = rc + hipMalloc(d_g, b2n)
      rc = rc + hipMalloc(d_h, 16_c_size_t)
      rc = rc + hipMalloc(d_p, bnp)
      rc = rc + hipMalloc(d_g2, 8_c_size_t)
      rc = rc + hipMalloc(d_age, 8_c_size_t)
      rc = rc + hipMalloc(d_rng, 8_c_size_t)
      rc = rc + hipMalloc(d_blk, bnp)
      if (rc .ne. 0) stop 'vmc (GPU): hipMalloc'
      w%nw = 1
      w%x = d_x
      w%g = d_g
      w%h = d_h
      w%props = d_p
      w%grad2_ph = d_g2
      w%age = d_age
      w%rng = d_rng
      w%blk = d_blk
c the walker and its stream to the device; the start evaluates it
      do i = 1, nptot
       hx(1,i) = x_old(1,i)
       hx(2,i) = x_old(2,i)
      enddo
      do i = 1, info%n_props
       hp(i) = p_old(i)
      enddo
      hg2(1) = grad2_ph
      hrng(1) = ((int(lrn(1),c_int64_t) * 4096 + lrn(2)) * 4096
     &          + lrn(3)) * 4096 + lrn(4)
      rc = hipMemcpy(d_x, c_loc(hx), b2n, hipMemcpyHostToDevice)
      rc = rc + hipMemcpy(d_p, c_loc(hp), bnp, hipMemcpyHostToDevice)
      rc = rc + hipMemcpy(d_g2, c_loc(hg2), 8_c_size_t,
     &                    hipMemcpyHostToDevice)
      rc = rc + hipMemcpy(d_rng, c_loc(hrng), 8_c_size_t,
     &                    hipMemcpyHostToDevice)
      if (rc .ne. 0) stop 'vmc (GPU): hipMemcpy'
      rc = qmc_set_delta(ctx, delta)
      rc = qmc_vmc_start(ctx, w, c_null_ptr)
      if (rc .ne. QMC_OK) then
       write(6,*) 'qmc_vmc_start: ', trim(qmc_error_text())
       stop 'vmc (GPU): start'
      endif
      do iblk = iblk0, nblk
       call averages(1, iblk, alg, uno)
       rc = hipMemcpy(d_blk, c_loc(hz), bnp, hipMemcpyHostToDevice)
       istp = 0
       do while (istp .lt. nstp)
        nchunk = min(nstp - istp, nskip - mod(istp, nskip))
        rc = qmc_vmc_sweep(ctx, w, nchunk, delta, mage, 0,
     &                     c_null_ptr, c_null_ptr, c_null_ptr)
        if (rc .ne. QMC_OK) then
         write(6,*) 'qmc_vmc_sweep: ', trim(qmc_error_text())
         stop 'vmc (GPU): sweep'
        endif
        istp = istp + nchunk
        if (mod(istp, nskip) .eq. 0) then
         call vmc_hip_walker_back
         call putconf(1)
        endif
       enddo
c averages(2) of every step: blk_av += sum of p_old, blk_norm += nstp
       call vmc_hip_walker_back
       rc = hipMemcpy(c_loc(hb), d_blk, bnp, hipMemcpyDeviceToHost)
       do i = 1, n_props
        psave(i) = p_old(i)
        p_old(i) = hb(i)
       enddo
       call averages(2, iblk, alg, uno)
       do i = 1, n_props
        p_old(i) = 0.d0
       enddo
       do istp = 2, nstp
        call averages(2, iblk, alg, uno)
       enddo
       do i = 1, n_props
        p_old(i) = psave(i)
       enddo
       call averages(3, iblk, alg, uno)
      enddo
      call vmc_hip_walker_back
      if (nstack .eq. 0) then
       call putconf(1)
      endif
      call write_conf
      rc = hipFree(d_x)
      rc = hipFree(d_g)
      rc = hipFree(d_h)
      rc = hipFree(d_p)
      rc = hipFree(d_g2)
      rc = hipFree(d_age)
      rc = hipFree(d_rng)
      rc = hipFree(d_blk)
      if (nitc .ne. 0) nitc = nitc0
      if (ncmass .ne. 0) ncmass = ncmass0
      if (nmstar .ne. 0) nmstar = nmstar0
      if (ntheta .ne. 0 .and. res_string .ne. '.') iblk0 = 1
      return
      contains
       subroutine vmc_hip_walker_back
c the device walker into COMMON (the state metropolis_test leaves)
        integer ip, it, rc2
        rc2 = hipMemcpy(c_loc(hx), d_x, b2n, hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hg), d_g, b2n,
     &                        hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hh), d_h, 16_c_size_t,
     &                        hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hp), d_p, bnp,
     &                        hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hg2), d_g2, 8_c_size_t,
     &                        hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hage), d_age, 8_c_size_t,
     &                        hipMemcpyDeviceToHost)
        rc2 = rc2 + hipMemcpy(c_loc(hrng), d_rng, 8_c_size_t,
     &                        hipMemcpyDeviceToHost)
        if (rc2 .ne. 0) stop 'vmc (GPU): hipMemcpy back'
        do ip = 1, nptot
         x_old(1,ip) = hx(1,ip)
         x_old(2,ip) = hx(2,ip)
         g_old(1,ip) = hg(1,ip)
         g_old(2,ip) = hg(2,ip)
        enddo
        do it = 1, ntypes
         h_old(it) = hh(it)
        enddo
        do ip = 1, n_props
         p_old(ip) = hp(ip)
        enddo
        s_old = 1.d0
        grad2_ph = hg2(1)
        age = hage(1)
        lrn(1) = int(iand(ishft(hrng(1), -36), 4095_c_int64_t))
        lrn(2) = int(iand(ishft(hrng(1), -24), 4095_c_int64_t))
        lrn(3) = int(iand(ishft(hrng(1), -12), 4095_c_int64_t))
        lrn(4) = int(iand(hrng(1), 4095_c_int64_t))
       end subroutine vmc_hip_walker_back
      end
