function [ccfmax, phi_max, psi_max, theta_max] = tom_oscar(vol, template, phi_start, phi_end, phi_incr, ...
        psi_start, psi_end, psi_incr, the_start, the_end, the_incr)
%TOM_OSCAR exhaustive unnormalised cross-correlation over an Euler grid on
%   the GPU: the signature and outputs of Filtrans/tom_oscar.m:1-82, with the
%   angle loop run by the tom_oscar_hip gateway (libtomhip).
%
%   The grid is phi_start:phi_incr:phi_end (outer), psi, then theta (inner),
%   as the reference's three nested loops; a voxel takes the angles of the
%   first orientation that strictly raises its correlation (-1000 and -1 where
%   none does). Deviation: the template is pasted at size/2+1-n/2 in every
%   dimension (the reference's y range reuses offsetx, Filtrans/tom_oscar.m:71).

angles = zeros(3, 0);
for phi = phi_start:phi_incr:phi_end
    for psi = psi_start:psi_incr:psi_end
        for the = the_start:the_incr:the_end
            angles(:, end + 1) = [phi; psi; the]; %#ok<AGROW>
        end
    end
end
[ccfmax, idx] = tom_oscar_hip(single(vol), single(template), single(angles));
ccfmax = double(ccfmax);
phi_max = int16(-ones(size(idx)));
psi_max = phi_max;
theta_max = phi_max;
hit = idx > 0;
phi_max(hit) = int16(angles(1, idx(hit)));
psi_max(hit) = int16(angles(2, idx(hit)));
theta_max(hit) = int16(angles(3, idx(hit)));
