function build_sgld_mex(repoRoot)
%BUILD_SGLD_MEX Compiles the MEX gateways of this folder against libsgldfea.
%   build_sgld_mex('/path/to/repo') after python sgldbench_amd/build.py has
%   produced sgldbench_amd/libsgldfea.so (gfx950).  The six stateless
%   gateways carry the names of the MEX functions they replace
%   (src/MEXfuncs/*.cpp), so the reference's .m files call them unchanged;
%   sgld_ctx_mex serves the shims (sgld_gpu_setup and the GPUsolve_ branches).
	here = fileparts(mfilename('fullpath'));
	inc = ['-I' fullfile(repoRoot, 'include')];
	lib = ['-L' fullfile(repoRoot, 'sgldbench_amd')];
	gws = {'Solving_KbyU_MatrixFree_mex', 'Solving_Restriction_MatrixFree_mex', ...
	       'Solving_Interpolation_MatrixFree_mex', 'Solving_AssembleCmptStencilFromFinestLevel', ...
	       'Solving_AssembleCmptStencilFromNonFinestLevel', 'TopOpti_CmptUnitCompliance_mex', 'sgld_ctx_mex'};
	for k = 1:numel(gws)
		mex('-largeArrayDims', inc, lib, '-lsgldfea', 'CXXFLAGS=$CXXFLAGS -std=c++17', ...
		    '-outdir', here, fullfile(here, [gws{k} '.cpp']));
	end
end
