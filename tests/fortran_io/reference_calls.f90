! Compile-only check of the drop-in module: the reference's own call lines
! for the five generic names, with the band arrays declared as wf_module
! keeps them (unk(ML_0:ML_1, MB, MBZ_0:MBZ_1, MSP_0:MSP_1), esp and res
! (MB, MBZ_0:MBZ_1, MSP_0:MSP_1)), against `use rsdft_gpu` in place of the
! reference's modules.  Call sites: scf_module.f90:231 (subspace_diag), :254
! (conjugate_gradient), :258 (gram_schmidt), :268 (esp_calc), main.f90:392
! (gram_schmidt) and cg_module.f90:189 (hamiltonian).
subroutine reference_calls(ML_0, ML_1, MB, Nband, MB_0, MB_1, ns, ne, nn)
  use rsdft_gpu
  implicit none
  integer, intent(in) :: ML_0, ML_1, MB, Nband, MB_0, MB_1, ns, ne, nn
  integer, parameter :: MBZ_0 = 1, MBZ_1 = 1, MSP_0 = 1, MSP_1 = 1
  real(8), allocatable :: unk(:, :, :, :), esp(:, :, :), res(:, :, :), hxk(:, :)
  integer :: k, s
  allocate (unk(ML_0:ML_1, MB, MBZ_0:MBZ_1, MSP_0:MSP_1))
  allocate (esp(MB, MBZ_0:MBZ_1, MSP_0:MSP_1), res(MB, MBZ_0:MBZ_1, MSP_0:MSP_1))
  allocate (hxk(ML_0:ML_1, nn))
  unk = 0d0; esp = 0d0; res = 0d0; hxk = 0d0
  do s = MSP_0, MSP_1
  do k = MBZ_0, MBZ_1
     call gram_schmidt(1,Nband,k,s)
  end do
  end do
  k = 1; s = 1
  call subspace_diag( k, s, unk(:,:,k,s), esp(:,k,s) )
  call conjugate_gradient(ML_0,ML_1,k,s,unk,esp,res)
  call gram_schmidt(1,Nband,k,s)
  call esp_calc( unk(:,MB_0:MB_1,k,s), esp(MB_0:MB_1,k,s), MB_0,k,s )
  call hamiltonian( unk(:,ns:ne,k,s), hxk(:,1:nn), ns,k,s )
end subroutine reference_calls
