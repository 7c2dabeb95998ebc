      program refeval
c Test-only harness around the reference Fortran (never shipped).
c It calls the reference's own routines and dumps their outputs so
c that the C restatement and the HIP library can be checked against
c them.  Linked with reptation.f + fake_mpi/mpif.f by build_ref.sh.
c
c Control file refeval.ctl (list-directed), in the deck directory:
c   mode                 1 = fixed-configuration evaluation
c                        2 = VMC move/accept replay
c                        3 = timing: mode 1 without output, repeated
c                            ncfg times over the first config block
c   ith jth              twist indices (used when the deck sets ntwists)
c   ncfg                 configurations (mode 1) or moves (mode 2)
c   delta                time step of the move (mode 2, 4); in mode 1
c                        a positive value sets the COMMON delta (sigma)
c                        read by the drift cutoff of decks with adrift
c   cfgfile              raw float64 positions, [ncfg][nptot][2]
c   outfile              raw output stream (layout below)
c
c Mode 4 only, line 7:
c   ntau ntauskip mdelta etrial mage mage_r every
c
c Mode 5 follows branch2 (reptation.f:2162-2271) on one rank: ncfg
c walkers in stack slots 1..ncfg with the weights of cfgfile (raw
c float64, ncfg values), the second generator seeded by input from the
c deck's seed file; slot i's first coordinate holds i, so that the
c redistribution (prov_dest) shows where each slot's configuration came
c from.  Output: the slot origins, nmult (weights clamped at wmult) and
c the second generator's limbs after the draws.
c
c Mode 1 follows read_conf (reptation.f:2452-2502) without the stack:
c x_new <- config; compute_properties(1) (reptation.f:2661).
c Mode 2 follows vmc (reptation.f:973-1017): one initial
c compute_properties + metropolis_test(1) as read_conf does, then
c move (reptation.f:4617) + metropolis_test (reptation.f:2622).
c Mode 4 follows rmc (reptation.f:1286-1313): read_path (reptation.f:
c 1766-1802, which reads the ntau initial beads from <x_file>.0 per
c type), then reptation_move (reptation.f:1897) +
c reptation_metropolis_test (reptation.f:1865) per move.  Every `every`
c moves the whole reptile
c state is dumped before the move.
      implicit none
      include 'ewald.h'
      integer mode,ncfg,icfg,ip,idim,ith0,jth0,i,it,k,seedv(4)
      integer istp,nk_out,nbuf,every,idelta,j,jstart,jnext,rtau
      integer rskip,rdelta
      integer*8 nrate,t0,t1
      parameter(nbuf=64)
      real*8 p,del,xb(2,mnp,nbuf),wnew,retrial,rmage,rmage_r,pb(64)
      real*8 wb(mstack)
      integer mb(mstack)
      character*256 cfgfile,outfile
      mytid=0
      nproc=1
      call input
      open(11,file='refeval.ctl',status='old')
      read(11,*) mode
      read(11,*) ith0,jth0
      read(11,*) ncfg
      read(11,*) del
      read(11,'(a)') cfgfile
      read(11,'(a)') outfile
      if(mode.eq.4)read(11,*) rtau,rskip,rdelta,retrial,rmage,rmage_r
     &                       ,every
      close(11)
      if(ntheta.ne.0)call twistg(ith0,jth0)
c the g-vector list in use (twistg or input's untwisted list)
      open(14,file='gvec.bin',access='stream',form='unformatted'
     &    ,status='replace')
      write(14) 2*nk+1
      write(14) ((gvec(idim,k),idim=1,2),k=1,2*nk+1)
      close(14)
      nk_out=0
      if(nrhok.ne.0)nk_out=nk
      open(12,file=cfgfile,access='stream',form='unformatted'
     &    ,status='old')
      open(13,file=outfile,access='stream',form='unformatted'
     &    ,status='replace')
      write(13) n_scalar_props,nptot,ntypes,nk_out,nrhok
      if(mode.eq.1)then
       if(del.gt.0.d0)then
        delta=del
        call sigma
       endif
       do icfg=1,ncfg
        read(12) ((x_new(idim,ip),idim=1,2),ip=1,nptot)
        call compute_properties(1)
        write(13) (p_new(i),i=1,n_scalar_props)
        write(13) ((g_new(idim,ip),idim=1,2),ip=1,nptot)
        write(13) (h_new(it),it=1,ntypes)
        write(13) grad2_ph
        do it=1,nrhok
         write(13) (rhok(k,it),k=1,2*nk)
        enddo
       enddo
      elseif(mode.eq.2)then
       delta=del
       call sigma
       mage=1.d30
       read(12) ((x_new(idim,ip),idim=1,2),ip=1,nptot)
       call compute_properties(1)
       p=1.d0
       call metropolis_test(p)
       do istp=1,ncfg
        call savern(seedv)
c per move: RNG state and walker before the move, then p, age after
c the test, proposed -log psi and local energy, kept -log psi, energy
        write(13) seedv
        write(13) ((x_old(idim,ip),idim=1,2),ip=1,nptot)
        call move(p)
        call metropolis_test(p)
        write(13) p,age,p_new(jltf),p_new(jetot),p_old(jltf)
     &           ,p_old(jetot)
       enddo
       write(13) ((x_old(idim,ip),idim=1,2),ip=1,nptot)
       call savern(seedv)
       write(13) seedv
      elseif(mode.eq.3)then
       do icfg=1,nbuf
        read(12) ((xb(idim,ip,icfg),idim=1,2),ip=1,nptot)
       enddo
       call system_clock(t0,nrate)
       do istp=1,ncfg
        icfg=mod(istp-1,nbuf)+1
        do ip=1,nptot
         x_new(1,ip)=xb(1,ip,icfg)
         x_new(2,ip)=xb(2,ip,icfg)
        enddo
        call compute_properties(1)
       enddo
       call system_clock(t1,nrate)
       write(6,'(a,i10,a,es16.8)') 'REFEVAL_TIMING evals ',ncfg
     &      ,' seconds ',dble(t1-t0)/dble(nrate)
       write(13) dble(t1-t0)/dble(nrate),p_new(jetot)
      elseif(mode.eq.4)then
       delta=del
       ntau=rtau
       ntauskip=rskip
       mdelta=rdelta
       etrial=retrial
       mage=rmage
       mage_r=rmage_r
       call sigma
       call read_path
       write(13) n_props_in_stack,n_buffer,n_props
       do istp=1,ncfg
        if(mod(istp-1,every).eq.0)then
c reptile state before the move (all n_buffer ring slots)
         call savern(seedv)
         write(13) seedv,jfirst,jlast
         write(13) age,grad2_ph
         write(13) (((x_stack(idim,ip,j),idim=1,2),ip=1,nptot)
     &             ,j=1,n_buffer)
         write(13) (((g_stack(idim,ip,j,iinc),idim=1,2),ip=1,nptot)
     &             ,j=1,n_buffer)
         write(13) ((p_stack(ip,j,iinc),ip=1,n_props_in_stack)
     &             ,j=1,n_buffer)
         write(13) (age_stack(j),j=1,n_buffer)
         write(13) ((w_stack(j,k),j=1,n_buffer),k=1,5)
         write(13) (s_stack(j),j=1,n_buffer)
         write(13) (p_p_old(i),i=1,n_props)
        endif
        call reptation_move(wnew)
c direction, beads attempted, path weight, Langevin p of each bead
        if(idir.eq.1)then
         jstart=jlast
        else
         jstart=jfirst
        endif
        do idelta=1,mdelta
         pb(idelta)=-1.d0
         if(idelta.le.ndelta)then
          jnext=mod(jstart-1+idir*idelta+n_buffer,n_buffer)+1
          pb(idelta)=p_stack(jacc,jnext,iinc)
         endif
        enddo
        write(13) idir,ndelta
        write(13) wnew,(pb(idelta),idelta=1,mdelta)
        call reptation_metropolis_test(wnew)
        write(13) age,p_p_old(jetot),p_p_old(jltf),p_p_old(jkin(0))
       enddo
       call savern(seedv)
       write(13) seedv,jfirst,jlast
      elseif(mode.eq.5)then
       do i=1,ncfg
        read(12) wb(i)
        x_stack(1,1,i)=dble(i)
       enddo
       getnext=1
       putnext=mod(ncfg,mstack)+1
       nstack=ncfg
       nconf=ncfg
       nmult=0
       call branch2(wb,mb)
       write(13) (nint(x_stack(1,1,i)),i=1,ncfg)
       call savern2(seedv)
       write(13) nmult,seedv
      endif
      close(12)
      close(13)
      stop
      end
