% Build the MATLAB gateways clODEmex, clODEfeaturesmex and clODEtrajectorymex
% from clode_mex.cpp against the native library libclode_hip.so.  Run from
% this folder after `python -m clode_amd.build`; put this folder on the
% MATLAB path so the reference's clODE, clODEfeatures and clODEtrajectory
% classes find the gateways by name.
here = fileparts(mfilename('fullpath'));
pkg = fileparts(here);                 % clode_amd/
repo = fileparts(pkg);
inc = ['-I' fullfile(repo, 'include')];
lib = ['-L' pkg];
rpath = ['LDFLAGS=$LDFLAGS -Wl,-rpath,' pkg];
names = {'clODEmex', 'clODEfeaturesmex', 'clODEtrajectorymex'};
for kind = 0:2
    mex('-R2018a', inc, lib, '-lclode_hip', rpath, sprintf('-DCLODE_MEX_KIND=%d', kind), ...
        fullfile(here, 'clode_mex.cpp'), '-output', fullfile(here, names{kind + 1}));
end
