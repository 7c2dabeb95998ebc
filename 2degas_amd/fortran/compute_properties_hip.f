      module qmc_hip_host
c The device context the Fortran forwarders share (created by the first
c compute_properties call from the deck state in COMMON) and the HIP
c runtime calls they use for device buffers.
      use iso_c_binding
      use qmc_hip_mod
      implicit none
      type(c_ptr), save :: ctx = c_null_ptr
      type(qmc_info), save :: info
      integer(c_int), parameter :: hipMemcpyHostToDevice = 1
      integer(c_int), parameter :: hipMemcpyDeviceToHost = 2
      interface
       integer(c_int) function hipMalloc(p, bytes)
     &   bind(C, name='hipMalloc')
        import :: c_int, c_ptr, c_size_t
        type(c_ptr) :: p
        integer(c_size_t), value :: bytes
       end function hipMalloc
       integer(c_int) function hipFree(p) bind(C, name='hipFree')
        import :: c_int, c_ptr
        type(c_ptr), value :: p
       end function hipFree
       integer(c_int) function hipMemcpy(dst, src, bytes, kind)
     &   bind(C, name='hipMemcpy')
        import :: c_int, c_ptr, c_size_t
        type(c_ptr), value :: dst, src
        integer(c_size_t), value :: bytes
        integer(c_int), value :: kind
       end function hipMemcpy
      end interface
      end module qmc_hip_host

      subroutine compute_properties(ifdist)
c Drop-in replacement of the reference's compute_properties
c (reptation.f:2661-2727).  Same name, same argument, same COMMON inputs
c and outputs: it reads x_new and the deck state left by `input`, and
c writes p_new(1..n_props), g_new, h_new, s_new, grad2_ph and rhok.  The
c evaluation runs on the GPU through lib2degas_hip.so (qmc_eval_batch with
c one walker).  Link it instead of the reference's compute_properties.
c
c The device context is created on the first call from the COMMON state
c (tables, k-vectors, box, particle types); the g-vector list is uploaded
c again whenever tabc/twistg change it.  The pair arrays of /c_dist/ are
c not filled (only gofr and ltf_bf2 read them, both outside this path).
      use iso_c_binding
      use qmc_hip_mod
      use qmc_hip_host
      implicit none
      include 'ewald.h'
      integer ifdist
      type(qmc_tables), save :: tab
      real(c_double), allocatable, target, save :: ut_c(:,:,:)
      integer(c_int32_t), allocatable, target, save :: ngrid_c(:)
      integer(c_int32_t), allocatable, target, save :: iexp_c(:)
      integer(c_int32_t), allocatable, target, save :: nke_c(:)
      real(c_double), allocatable, target, save :: tail_c(:)
      real(c_double), allocatable, target, save :: kvec_c(:,:)
      real(c_double), allocatable, save :: gsent(:,:)
      real(c_double), allocatable, save :: xw(:,:), gw(:,:), pw(:)
      real(c_double) :: hw(2), g2w(1)
      integer(c_int32_t) :: stw(1)
      integer :: i, j, k, it, jt, ntab, ng, rc
      logical :: changed
c first call: copy the deck state into the device context
      if (.not. c_associated(ctx)) then
       ntab = 0
       do it = 1, ntypes
        do jt = 1, ntypes
         ntab = max(ntab, iu2table(it,jt,iinc), iu3table(it,jt,iinc),
     &              iubtable(it,jt,iinc), iv2table(it,jt,iinc))
        enddo
       enddo
       allocate(ut_c(0:mgrid,4,max(ntab,1)), ngrid_c(max(ntab,1)),
     &          iexp_c(max(ntab,1)), nke_c(max(ntab,1)),
     &          tail_c(max(ntab,1)), kvec_c(2,max(nk,1)))
       do k = 1, ntab
        ut_c(:,:,k) = ut(:,:,k)
        ngrid_c(k) = ngrid(k)
        iexp_c(k) = iexp(k)
        nke_c(k) = nk_ewald(k)
        tail_c(k) = tail(k)
       enddo
       do k = 1, nk
        kvec_c(1,k) = kvec(1,k)
        kvec_c(2,k) = kvec(2,k)
       enddo
       tab%ndim = ndim
       tab%ntypes = ntypes
       tab%np = 0
       tab%hbs2m = 0.d0
       do it = 1, ntypes
        tab%np(it) = np(it)
        tab%hbs2m(it) = hbs2m(it)
       enddo
       tab%el(1) = el(1)
       tab%el(2) = el(2)
       tab%drt = drt
       tab%ntab = ntab
       tab%ldut = mgrid + 1
       tab%ut = c_loc(ut_c)
       tab%ngrid = c_loc(ngrid_c)
       tab%tail = c_loc(tail_c)
       tab%iexp = c_loc(iexp_c)
       tab%nk_ewald = c_loc(nke_c)
       tab%iu2 = 0
       tab%iu3 = 0
       tab%iub = 0
       tab%iv2 = 0
       tab%v2value = 1.d0
       do it = 1, ntypes
        do jt = 1, ntypes
         k = (it-1)*2 + jt
         tab%iu2(k) = iu2table(it,jt,iinc)
         tab%iu3(k) = iu3table(it,jt,iinc)
         tab%iub(k) = iubtable(it,jt,iinc)
         tab%iv2(k) = iv2table(it,jt,iinc)
         tab%v2value(k) = v2value(it,jt,iinc)
        enddo
       enddo
       tab%v0 = v0
       tab%nk = nk
       tab%kvec = c_loc(kvec_c)
       tab%irhok = 0
       do it = 1, ntypes
        tab%irhok(it) = irhok(it)
       enddo
       tab%nrhok = nrhok
       tab%ntheta = ntheta
       tab%adrift = adrift
       tab%delta = delta
       rc = qmc_ctx_create(tab, 0, ctx)
       if (rc .ne. QMC_OK) then
        write(6,*) 'qmc_ctx_create: ', trim(qmc_error_text())
        stop 'compute_properties (GPU): context'
       endif
       rc = qmc_ctx_info(ctx, info)
       allocate(xw(2,nptot), gw(2,nptot), pw(info%n_props))
       allocate(gsent(2, 2*nk+1))
       gsent = huge(1.d0)
      endif
c g-vectors: re-send when tabc/twistg changed them
      ng = 2*nk + 1
      changed = .false.
      do k = 1, ng
       if (gsent(1,k) .ne. gvec(1,k) .or. gsent(2,k) .ne. gvec(2,k))
     &   changed = .true.
      enddo
      if (changed) then
       do k = 1, ng
        gsent(1,k) = gvec(1,k)
        gsent(2,k) = gvec(2,k)
       enddo
       rc = qmc_set_twist(ctx, gsent, ng)
       if (rc .ne. QMC_OK) stop 'compute_properties (GPU): g-vectors'
      endif
c the drift cutoff reads the current COMMON delta (reptation.f:2755)
      rc = qmc_set_delta(ctx, delta)
c one walker through the C-ABI
      do i = 1, nptot
       xw(1,i) = x_new(1,i)
       xw(2,i) = x_new(2,i)
      enddo
      do i = 1, info%n_props
       pw(i) = p_new(i)
      enddo
      g2w(1) = grad2_ph
      rc = qmc_eval_batch(ctx, 1, xw, pw, gw, hw, g2w, stw)
      if (rc .eq. QMC_ERR_SINGULAR)
     &  stop 'slater_bckflw_orbitals (1): info.ne.0'
      if (rc .ne. QMC_OK) then
       write(6,*) 'qmc_eval_batch: ', trim(qmc_error_text())
       stop 'compute_properties (GPU): evaluation'
      endif
      do i = 1, info%n_props
       p_new(i) = pw(i)
      enddo
      do i = 1, nptot
       g_new(1,i) = gw(1,i)
       g_new(2,i) = gw(2,i)
      enddo
      do it = 1, ntypes
       h_new(it) = hw(it)
      enddo
      s_new = 1.d0
      grad2_ph = g2w(1)
c rhok as distances leaves it (reptation.f:4224-4243)
      j = info%jrhok
      do it = 1, ntypes
       if (irhok(it) .ne. 0) then
        do k = 1, 2*nk
         rhok(k, irhok(it)) = pw(j)
         j = j + 1
        enddo
       endif
      enddo
      return
      end
