function rTilde = Solving_Vcycle(r)
%SOLVING_VCYCLE Shim over src/Solving_Vcycle.m:1-52 (rename the reference file
%   Solving_Vcycle_cpu.m): the additive multigrid preconditioner, on the
%   device context when GPUsolve_ is true.
	global GPUsolve_ sgldCtx_
	if ~isempty(GPUsolve_) && GPUsolve_
		rTilde = sgld_ctx_mex('vcycle', sgldCtx_, r);
	else
		rTilde = Solving_Vcycle_cpu(r);
	end
end
