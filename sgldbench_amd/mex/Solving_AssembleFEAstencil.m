function Solving_AssembleFEAstencil()
%SOLVING_ASSEMBLEFEASTENCIL Shim over src/Solving_AssembleFEAstencil.m:1-138
%   (rename the reference file Solving_AssembleFEAstencil_cpu.m).  With
%   GPUsolve_ true the Galerkin coarse operators, Jacobi diagonals and the
%   coarsest Cholesky factor are built on the device from
%   meshHierarchy_(1).eleModulus.
	global meshHierarchy_ weightFactorJacobi_ GPUsolve_ sgldCtx_
	if ~isempty(GPUsolve_) && GPUsolve_
		sgld_ctx_mex('set_modulus', sgldCtx_, meshHierarchy_(1).eleModulus(:));
		sgld_ctx_mex('assemble', sgldCtx_, weightFactorJacobi_);
	else
		Solving_AssembleFEAstencil_cpu();
	end
end
