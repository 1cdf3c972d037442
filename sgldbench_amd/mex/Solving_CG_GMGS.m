function varargout = Solving_CG_GMGS(printP)
%SOLVING_CG_GMGS Shim over src/Solving_CG_GMGS.m:1-9 (rename the reference
%   file Solving_CG_GMGS_cpu.m).  With GPUsolve_ true the whole MG-PCG solve
%   F_ -> U_ (U_ in: warm start) runs in the device context sgldCtx_;
%   GPUprecision_ 'mixed' selects the FP32-inner solve with FP64 reliable
%   updates, anything else the FP64 reference recurrence.
	global U_ F_ tol_ maxIT_ GPUsolve_ GPUprecision_ sgldCtx_
	if ~isempty(GPUsolve_) && GPUsolve_
		prec = 'fp64';
		if ~isempty(GPUprecision_), prec = GPUprecision_; end
		[U_, its, relres] = sgld_ctx_mex('pcg', sgldCtx_, F_, U_, tol_, maxIT_, prec);
		if strcmp(printP, 'printP_ON')
			fprintf('GPU MG-PCG (%s): %d iterations, relative residual %.3e\n', prec, its, relres);
		end
	else
		its = Solving_CG_GMGS_cpu(printP);
	end
	if 1==nargout, varargout{1} = its; end
end
