function productMV = Solving_KbyU_MatrixFree(uVec)
%SOLVING_KBYU_MATRIXFREE Shim over src/Solving_KbyU_MatrixFree.m:1-41 (rename
%   the reference file Solving_KbyU_MatrixFree_cpu.m): K u with fixed DOFs
%   zeroed, on the device context when GPUsolve_ is true.
	global GPUsolve_ sgldCtx_
	if ~isempty(GPUsolve_) && GPUsolve_
		productMV = sgld_ctx_mex('kbyu', sgldCtx_, uVec);
	else
		productMV = Solving_KbyU_MatrixFree_cpu(uVec);
	end
end
