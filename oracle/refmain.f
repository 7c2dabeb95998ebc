      program refmain
c Test-only serial entry point of the reference driver: the same call
c sequence as the reference program entry (main.f:16-21), built with
c the reference's serial MPI stubs (fake_mpi/mpif.f).  Used to time
c the reference CPU path and to run its own VMC driver end to end.
      implicit none
      include 'mpif.h'
      include 'ewald.h'
      integer jrc
      call MPI_INIT(jrc)
      call MPI_COMM_SIZE(MPI_COMM_WORLD,nproc,jrc)
      call MPI_COMM_RANK(MPI_COMM_WORLD,mytid,jrc)
      call input
      call sonaseppia
      call MPI_FINALIZE(jrc)
      stop
      end
