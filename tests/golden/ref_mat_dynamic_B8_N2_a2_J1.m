function [J, N, B, bl_scoeff, alpha, scoeff, wcoeff] = ref_mat_dynamic_B8_N2_a2_J1
% function [J, N, B, bl_scoeff, alpha, scoeff, wcoeff] = ref_mat_dynamic_B8_N2_a2_J1
%
%  Scale discretised wavelet coefficients created by S2DW code
%  S2DW code written by Jason McEwen (mcewen@mrao.cam.ac.uk)
%
%  This file written by: Fortran
%  Comment: fits test
%
%  Variables:
%   J: Maximum analysis scale depth.
%   B: Harmonic band limit.
%   N: Azimuthal band limit.
%   bl_scoeff: Upper band limit for scaling coefficients.
%   alpha: Basis dilation factor.
%   scoeff(1:bl_scoeff,1:bl_scoeff): Scaling coefficients.
%   wcoeff{1:J+1}(aa,bb,gg): Wavelet coefficients (ranges of aa, bb and gg, depend on scale).

% Size parameters
J         =                           1;
N         =                           2;
B         =                           8;
bl_scoeff =                           4;
alpha     =  0.20000000000000000000E+01;

% Wavelet coefficient array sizes
wcoeff_size_aa(    1) =              15;
wcoeff_size_bb(    1) =              16;
wcoeff_size_gg(    1) =               2;
wcoeff_size_aa(    2) =              15;
wcoeff_size_bb(    2) =              16;
wcoeff_size_gg(    2) =               2;

% Scaling coefficients
scoeff_array = load('ref_mat_dynamic_B8_N2_a2_J1_scoeff.dat');
n = length(scoeff_array) / 2;
scoeff_array = scoeff_array(1:n) + i*scoeff_array(n+1:end);
scoeff = zeros(bl_scoeff,bl_scoeff);
ind_start = 1;
for el = 0:bl_scoeff-1
    ind_end = ind_start - 1 + el + 1;
    scoeff(el+1, 1:el+1) = scoeff_array(ind_start:ind_end).';
    ind_start = ind_end + 1;
end

% Wavelet coefficients
wcoeff_array = load('ref_mat_dynamic_B8_N2_a2_J1_wcoeff.dat');
ind_start = 1;
for jj = 1:J+1
    ind_end = ind_start - 1 + wcoeff_size_aa(jj) * wcoeff_size_bb(jj) * wcoeff_size_gg(jj);
    wcoeff{jj} = reshape(wcoeff_array(ind_start:ind_end),wcoeff_size_aa(jj), wcoeff_size_bb(jj), wcoeff_size_gg(jj));
    ind_start = ind_end+1;
end

