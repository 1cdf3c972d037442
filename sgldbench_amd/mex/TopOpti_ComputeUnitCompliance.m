function ceList = TopOpti_ComputeUnitCompliance()
%TOPOPTI_COMPUTEUNITCOMPLIANCE Shim over src/TopOpti_ComputeUnitCompliance.m:1-21
%   (rename the reference file TopOpti_ComputeUnitCompliance_cpu.m):
%   ce_e = u_e' Ke u_e for every element, on the device context when
%   GPUsolve_ is true.
	global U_ GPUsolve_ sgldCtx_
	if ~isempty(GPUsolve_) && GPUsolve_
		ceList = sgld_ctx_mex('unit_compliance', sgldCtx_, U_);
	else
		ceList = TopOpti_ComputeUnitCompliance_cpu();
	end
end
