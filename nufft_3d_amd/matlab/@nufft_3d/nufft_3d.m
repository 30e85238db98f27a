classdef nufft_3d < handle
    % MI355X (HIP) drop-in for the @nufft_3d class: the same constructor and
    % methods (nufft_3d.m:54, fNUFT.m:2, aNUFT.m:2, iNUFT.m:2, sake.m:1,
    % pruno.m:2, discrep.m:2), computed by libnufft3d through the MEX
    % gateway nufft3d_mex (nufft3d_mex.c beside this folder).
    %
    %   obj = nufft_3d(om,N,'J',4,'u',2,'radial',1,'gpu',1);
    %   k   = obj.fNUFT(x);        % x [N1 N2 N3 nc] -> k [M nc]
    %   x   = obj.aNUFT(k);
    %   im  = obj.iNUFT(raw,maxit,damp,W,constraint,lambda);
    %
    % Differences from the reference class, all deliberate:
    %  - a handle class: copies share one device-resident plan and delete
    %    frees it (the reference is a value class holding H and HT);
    %  - gpu=0 (MATLAB's CPU sparse path) is not offered; gpu=1 computes in
    %    single precision, gpu=2 in double (the library's FP64 mode);
    %  - arrays cross as host arrays (MATLAB gpuArray is CUDA only): inputs
    %    are gathered and cast, outputs come back as host arrays;
    %  - pruno returns K [N1 N2 N3 nc nv] as the reference does, built here
    %    from the library's tap-domain kernels, and also nv, the singular
    %    values and the noise level as further outputs; no plot is drawn;
    %  - the 'parallel-imaging-sake' branch returns the image ifft3(ksp)
    %    (iNUFT.m:195 returns SAKE's k-space as the image).

    properties (SetAccess = private)
        J(1,1) double       = 4          % kernel width (nufft_3d.m:27)
        u(1,1) double       = 2          % oversampling (nufft_3d.m:28)
        N(1,3) double = zeros(1,3)       % image size
        alpha(1,1) double   = 0          % Kaiser-Bessel alpha (0 = Beatty)
        radial(1,1) logical = 1          % radial kernel
        gpu(1,1) double     = 1          % 1 = single, 2 = double
        K(1,3) double = zeros(1,3)       % oversampled size
        nnz(1,1) double     = 0          % samples in bounds
        M(1,1) double       = 0          % samples passed (rows of H)
    end

    properties (Dependent)
        d                                % density weighting (density.m)
        U                                % deapodization matrix (deap.m)
    end

    properties (Access = private)
        h = uint64(0)                    % device-resident plan
        dUser = []
    end

    methods
        %% constructor (nufft_3d.m:54-127)
        function obj = nufft_3d(om,N,varargin)
            if ~isnumeric(om) || ~isreal(om) || isempty(om) || size(om,1)~=3
                error('om must be an array with leading dimension of 3');
            end
            om = reshape(double(gather(om)),3,[]);
            if ~exist('N','var') || isempty(N)
                N = 2*ceil(max(abs(om),[],2));
                warning('N argument not supplied - using [%i %i %i]',N);
            end
            for k = 1:2:numel(varargin)                  % nufft_3d.m:71-76
                if k==numel(varargin) || ~ischar(varargin{k})
                    error('''varargin'' must be option/value pairs.');
                end
                switch varargin{k}
                    case 'J',      obj.J = varargin{k+1};
                    case 'u',      obj.u = varargin{k+1};
                    case 'alpha',  obj.alpha = varargin{k+1};
                    case 'radial', obj.radial = varargin{k+1};
                    case 'gpu',    obj.gpu = varargin{k+1};
                    case 'd',      obj.dUser = double(gather(varargin{k+1}));
                    otherwise,     error('''%s'' is not a valid property.',varargin{k});
                end
            end
            if obj.gpu==0
                error('gpu=0 (the MATLAB CPU path) is not provided by the HIP library');
            end
            obj.h = nufft3d_mex('create',om,double(N),obj.J,obj.u,obj.alpha, ...
                                double(obj.radial),obj.gpu,obj.dUser(:));
            [obj.K,obj.N,obj.J,obj.u,obj.alpha,r,obj.gpu,obj.nnz] = nufft3d_mex('params',obj.h);
            obj.radial = logical(r);
            obj.M = size(om,2);
        end

        function delete(obj)
            if obj.h ~= 0
                nufft3d_mex('destroy',obj.h);
                obj.h = uint64(0);
            end
        end

        %% forward / adjoint (fNUFT.m:2-9, aNUFT.m:2-9)
        function k = fNUFT(obj,x)
            k = nufft3d_mex('fNUFT',obj.h,obj.cast(x),obj.M);
        end

        function x = aNUFT(obj,k)
            x = nufft3d_mex('aNUFT',obj.h,reshape(obj.cast(k),obj.M,[]));
        end

        function y = iprojection(obj,x,damp)             % nufft_3d.m:305-313
            if ~exist('damp','var') || isempty(damp), damp = 0; end
            y = nufft3d_mex('iprojection',obj.h,obj.cast(x),damp);
            y = reshape(y,prod(obj.N),[]);
        end

        %% inverse (iNUFT.m:2-227)
        function im = iNUFT(obj,raw,maxit,damp,W,constraint,lambda)
            nrow = obj.M;
            if size(raw,1)==nrow
                nc = size(raw,2);
            else
                nr = size(raw,1); ny = size(raw,2);
                if nr*ny ~= nrow
                    error('raw data leading dimension(s) must be length %i (not %ix%i)',nrow,nr,ny);
                end
                nc = size(raw,3);
            end
            raw = reshape(raw,nrow,nc,[]);
            if size(raw,3)>1
                error('raw data size [%s] seems to have too many dimensions',num2str(size(raw)));
            end
            if ~exist('maxit','var') || isempty(maxit)
                maxit = 20;
            else
                validateattributes(maxit,{'numeric'},{'scalar','finite','integer','nonnegative'},'','maxit');
            end
            if ~exist('damp','var') || isempty(damp)
                damp = 0;
            else
                validateattributes(damp,{'numeric'},{'scalar','finite','nonnegative'},'','damp');
            end
            if ~exist('W','var') || isscalar(W) || isempty(W)
                W = [];                                   % density only (iNUFT.m:55-56)
            elseif maxit<=1
                warning('weighting is only effective when maxit>1');
            end
            if ~exist('constraint','var') || isempty(constraint)
                constraint = '';
            elseif maxit<=1
                warning('phase constraint is only effective when maxit>1');
                constraint = '';
            else
                switch constraint
                    case 'phase-constraint';
                    case 'compressed-sensing';
                    case 'parallel-imaging-sake'; if nc==1; error('sake-low-rank requires multiple coils'); end
                    case 'parallel-imaging-pruno'; if nc==1; error('parallel-imaging requires multiple coils'); end
                    otherwise; error('unknown constraint');
                end
                if ~exist('lambda','var') || isempty(lambda)
                    error('lambda must be supplied with %s',constraint);
                end
                validateattributes(lambda,{'numeric'},{'scalar','finite','nonnegative'},'','lambda');
            end
            if ~exist('lambda','var') || isempty(lambda), lambda = 0; end
            im = nufft3d_mex('iNUFT',obj.h,obj.cast(raw),maxit,damp,double(gather(W(:))), ...
                             constraint,lambda);
        end

        %% SAKE (sake.m:1-171): options as name/value pairs
        function ksp = sake(obj,data,varargin)
            opts = struct();
            valid = {'width','radial','loraks','tol','maxit','noise','center','W','damp','sparsity'};
            for k = 1:2:numel(varargin)
                if k==numel(varargin) || ~ischar(varargin{k})
                    error('''varargin'' must be option/value pairs.');
                end
                if ~any(strcmp(varargin{k},valid))
                    warning('''%s'' is not a valid option.',varargin{k});
                end
                opts.(varargin{k}) = varargin{k+1};
            end
            if size(data,1)~=obj.M || size(data,2)<2 || ~isfloat(data)
                error('''data'' must be a float array [npts nc].');
            end
            if isfield(opts,'center') && ~isempty(opts.center), opts.center = double(opts.center); end
            ksp = nufft3d_mex('sake',obj.h,obj.cast(data),opts);
        end

        %% PRUNO (pruno.m:2-143): image-space null kernels, null-space
        %% size, spectrum and noise
        function [K,nv,sv,noise] = pruno(obj,data,varargin)
            opts = struct('width',5,'radial',1);
            for k = 1:2:numel(varargin)
                if k==numel(varargin) || ~ischar(varargin{k})
                    error('''varargin'' must be option/value pairs.');
                end
                opts.(varargin{k}) = varargin{k+1};
            end
            [nv,sv,noise,Kt] = nufft3d_mex('pruno',obj.h,obj.cast(data),opts);
            % kernel mask (pruno.m:56-65): x fastest, z = 0 for 2d data
            h = fix(opts.width/2);
            if obj.N(3)==1
                [x,y] = ndgrid(-h:h); z = zeros(size(x));
            else
                [x,y,z] = ndgrid(-h:h);
            end
            if opts.radial
                mask = sqrt(x.^2+y.^2+z.^2)<=opts.width/2;
            else
                mask = abs(x)<=opts.width/2 & abs(y)<=opts.width/2 & abs(z)<=opts.width/2;
            end
            % taps into the mask, then the zero-padded inverse DFT times N
            % per axis (pruno.m:124-143); Kt rows are coil fastest
            nc = size(Kt,1)/nnz(mask);
            K = zeros(numel(mask),nc,nv);
            K(mask,:,:) = permute(reshape(Kt,nc,[],nv),[2 1 3]);
            if obj.N(3)==1
                K = reshape(K,[size(mask) 1 nc nv]);
            else
                K = reshape(K,[size(mask) nc nv]);
            end
            K = ifft(K,obj.N(1),1) * obj.N(1);
            K = ifft(K,obj.N(2),2) * obj.N(2);
            K = ifft(K,obj.N(3),3) * obj.N(3);
        end

        %% discrepancy principle (discrep.m:2-69)
        function damp = discrep(obj,data,stddev)
            nrow = obj.M;
            if size(data,1)==nrow
                nc = size(data,2);
            else
                nr = size(data,1); ny = size(data,2);
                if nr*ny ~= nrow
                    error('data leading dimension(s) must be length %i (not %ix%i)',nrow,nr,ny);
                end
                nc = size(data,3);
            end
            data = reshape(data,nrow,nc,[]);
            if size(data,3)>1
                error('data size [%s] seems to have too many dimensions',num2str(size(data)));
            end
            if exist('stddev','var')
                validateattributes(stddev,{'numeric'},{'scalar','finite','nonnegative'},'','stddev');
                damp = nufft3d_mex('discrep',obj.h,obj.cast(data),stddev);
            else
                damp = nufft3d_mex('discrep',obj.h,obj.cast(data),[]);
            end
        end

        %% properties computed by the library
        function v = get.d(obj)
            v = nufft3d_mex('density',obj.h,obj.M);
        end

        function v = get.U(obj)
            v = nufft3d_mex('deap',obj.h);
            if obj.gpu==1, v = single(v); end
        end
    end

    methods (Access = private)
        function x = cast(obj,x)
            x = gather(x);
            if obj.gpu==1
                x = complex(single(x));
            else
                x = complex(double(x));
            end
        end
    end
end
